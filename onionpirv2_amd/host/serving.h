// Continuous-batching serving loop over several resident PIR databases.
//
// The reference's PirServer answers one query per make_query call
// (src/server.cpp:895-947) and has no serving layer. This loop is the system
// form on top of the same answer path: requests from many clients arrive at
// any time for any of several databases. Each database has `lanes` answer
// lanes -- its own context and database handle, and for every further lane a
// context clone with a view of the same resident database (opir_ctx_clone,
// opir_db_view) -- each with its own worker thread, HIP streams and
// workspace. A free lane takes whatever is queued for its database (up to the
// admitted batch size) as one fused batch (opir_answer_batch); while it runs,
// requests that arrive are taken by the next free lane, so consecutive batches
// of one database overlap on the device (one batch's expansion beside the
// previous one's scan and MUX tree), and batches of different databases
// overlap too. Batch sizes are admitted against an HBM budget for client keys
// and the lanes' workspaces (expanded queries, selectors, first-dimension
// candidates). The plan is an equal share per database, reserved up front.
// With shrink_after_s >= 0 it is re-planned as the load moves: a lane idle for
// shrink_after_s gives its workspace back down to one query's
// (opir_db_trim); a lane whose queue holds more than it may take grows into
// budget no lane holds; a lane with work that holds less than its share and
// cannot grow to it flags a reclaim, and lanes above their share trim back to
// it at their next batch boundary (and do not grow past it while the flag
// stands). Each grow or trim reallocates a workspace, which waits for the
// lane's queued work.
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "onionpir_host.h"

namespace onionpir {

struct ServeConfig {
  size_t max_batch = 64;      // upper bound per database and batch
  std::vector<size_t> max_batch_db;  // optional tighter bound per database (empty: max_batch for all)
  size_t hbm_budget = 0;      // bytes for keys + workspaces; 0 = bounded by max_batch only
  size_t keyed_clients = 0;   // clients per context counted against the budget; 0 = the keys resident at start
  size_t lanes = 2;           // concurrent answer lanes per database (1 = one batch at a time)
  double shrink_after_s = -1;  // re-planning: idle time after which a lane trims to one query; < 0: fixed plan
};

struct ServeResult {
  uint64_t ticket = 0;
  size_t db = 0, client_id = 0;
  std::vector<u64> resp;      // [c0 N][c1 N] mod small_q; empty when error is set
  double t_submit = 0, t_start = 0, t_done = 0;  // seconds since the loop started
  size_t batch = 0;           // size of the batch that answered it
  std::string error;
};

// Largest admitted batch per database: an equal share of (budget - fixed)
// per database, and the largest B <= max_batch whose workspace ws(d, B) fits
// that share (ws is non-decreasing in B). budget == 0 admits max_batch.
// Throws std::invalid_argument if some database cannot take even B = 1.
std::vector<size_t> plan_batch_limits(size_t ndb, const std::function<size_t(size_t, size_t)> &ws, size_t fixed,
                                      size_t budget, size_t max_batch);

class ServingLoop {
 public:
  ServingLoop(std::vector<PirServer *> dbs, const ServeConfig &cfg);
  ~ServingLoop();
  ServingLoop(const ServingLoop &) = delete;
  ServingLoop &operator=(const ServingLoop &) = delete;

  // Thread-safe. query: [c0 K*N][c1 K*N] coefficient form (K RNS limbs).
  uint64_t submit(size_t db, size_t client_id, std::vector<u64> query);
  // Moves completed results into out; returns how many.
  size_t poll(std::vector<ServeResult> &out);
  // Blocks until every submitted request has completed.
  void wait_idle();
  void stop();
  // The equal-share plan every lane starts from.
  const std::vector<size_t> &batch_limits() const { return limits_; }
  // Re-planning: the batch size lane `lane` of database `db` may take now,
  // the largest one any lane of `db` has held, and how often lanes grew and
  // trimmed their workspaces.
  size_t reserved_batch(size_t db, size_t lane) const;
  size_t peak_batch(size_t db) const;
  size_t grows() const;
  size_t trims() const;
  double now() const;

 private:
  struct Req {
    uint64_t ticket;
    size_t client_id;
    std::vector<u64> query;
    double t_submit;
  };
  struct DbLane {              // one database: its queue and its answer lanes
    PirServer *srv = nullptr;
    std::deque<Req> queue;
    struct Lane {
      opir_ctx *ctx = nullptr;
      opir_db *db = nullptr;
      bool clone = false;          // ctx / db are a clone and a view owned by the loop
      std::mutex *ctx_mu = nullptr;  // serialises lanes that share one context
      std::thread th;
      size_t res_B = 0, res_bytes = 0;  // reserved batch and its workspace bytes (under bmu_)
      bool reclaim = false;             // holds less than its share while it has work (under bmu_)
    };
    std::vector<Lane> lanes;
    size_t peak_B = 0;         // under bmu_
  };
  void worker(size_t d, size_t lane);
  size_t lane_ws(size_t d, size_t B) const;  // one lane's workspace bytes for B queries
  void try_grow(size_t d, size_t li, size_t want);
  void trim_to(size_t d, size_t li, size_t B);
  bool others_reclaim(size_t d, size_t li) const;  // under bmu_

  std::vector<std::unique_ptr<DbLane>> lanes_;
  std::vector<std::unique_ptr<std::mutex>> ctx_mus_;
  std::vector<size_t> limits_;
  size_t budget_ = 0, fixed_ = 0, reserved_total_ = 0, grows_ = 0, trims_ = 0;
  std::vector<size_t> caps_;   // largest batch per database
  double shrink_after_ = -1;
  mutable std::mutex bmu_;     // budget accounting: reserved_total_, lane res_*, peaks, counters
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<ServeResult> done_;
  uint64_t next_ticket_ = 0, submitted_ = 0, completed_ = 0;
  bool stop_ = false;
  std::chrono::steady_clock::time_point t0_;
};

// Open-loop load: Poisson arrivals at rate lambda (1/s) for `duration`
// seconds, each request a uniformly drawn entry of the pool (db, client,
// query). Returns every result (with the pool index it came from) once all
// have completed.
struct PoolEntry {
  size_t db, client_id;
  std::vector<u64> query;
};
struct LoadRun {
  std::vector<ServeResult> results;
  std::vector<size_t> pool_index;  // per result
  double offered_qps = 0, achieved_qps = 0, span_s = 0;
};
LoadRun run_poisson(ServingLoop &loop, const std::vector<PoolEntry> &pool, double lambda, double duration, u64 seed);

}  // namespace onionpir
