// Request loop for SpGEMM serving: a plain mode (one row-wise call per
// request, in arrival order) and a continuously batched mode over a set of
// batch engines.
//
// Batched mode, per iteration of the loop:
//   1. completions: every engine whose batch finished releases its requests
//      (their finish time is the time the loop observes the completion);
//   2. arrivals: requests whose arrival time has passed join the queue;
//   3. admission: each idle engine takes queued requests in arrival order
//      while its pools hold them (rows, operand nonzeros, and the upper bound
//      of the output entries) and the batch stays under the products cap,
//      then launches its graph.  Requests admitted
//      while other engines run overlap with them on the device (one stream
//      per engine).  A request no engine can ever hold runs through the
//      plain path.
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "common.hpp"

using namespace spws;

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

void wait_until(Clock::time_point t0, double t) {
  for (;;) {
    const double now = since(t0);
    if (now >= t) return;
    if (t - now > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

int run_plain(spws_ctx* ctx, const spws_request& q) {
  int64_t nnz;
  int32_t *p, *c;
  double* v;
  return spws_spgemm_rowwise(ctx, q.m, q.k, q.n, q.A_pos, q.A_crd, q.A_vals, q.B_pos, q.B_crd, q.B_vals, &nnz, &p, &c,
                             &v);
}

}  // namespace

extern "C" {

int spws_serve(spws_ctx* ctx, spws_batch** engines, int32_t nengines, int32_t nreq, const spws_request* reqs,
               const spws_serve_options* opt, double* finish, int32_t* batch_of, spws_serve_stats* stats) {
  SPWS_REQUIRE(ctx && reqs && opt && finish && nreq >= 0, SPWS_ERR_INVALID, "null argument");
  for (int32_t i = 1; i < nreq; ++i)
    SPWS_REQUIRE(reqs[i].arrival >= reqs[i - 1].arrival, SPWS_ERR_INVALID, "requests must be sorted by arrival");
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  spws_serve_stats st{};
  const Clock::time_point t0 = Clock::now();
  if (opt->mode == 0) {
    for (int32_t r = 0; r < nreq; ++r) {
      wait_until(t0, reqs[r].arrival);
      if ((rc = run_plain(ctx, reqs[r]))) return rc;
      finish[r] = since(t0);
      if (batch_of) batch_of[r] = -1;
    }
    st.wall_seconds = since(t0);
    if (stats) *stats = st;
    return SPWS_OK;
  }
  SPWS_REQUIRE(engines && nengines >= 1, SPWS_ERR_INVALID, "batched mode needs at least one engine");
  const int32_t cap_req = opt->max_batch_requests > 0 ? opt->max_batch_requests : INT32_MAX;
  const int64_t cap_prod = opt->max_batch_products > 0 ? opt->max_batch_products : INT64_MAX;
  std::vector<std::vector<int32_t>> owned((size_t)nengines);
  std::vector<bool> busy((size_t)nengines, false);
  std::deque<int32_t> queue;
  int32_t next = 0, done = 0;
  int64_t batch_id = 0;
  while (done < nreq) {
    bool progress = false;
    for (int32_t e = 0; e < nengines; ++e) {
      if (!busy[(size_t)e]) continue;
      int fin = 0;
      if ((rc = spws_batch_poll(engines[e], &fin))) return rc;
      if (fin) {
        const double t = since(t0);
        for (int32_t r : owned[(size_t)e]) finish[r] = t;
        done += (int32_t)owned[(size_t)e].size();
        owned[(size_t)e].clear();
        busy[(size_t)e] = false;
        progress = true;
      }
    }
    while (next < nreq && reqs[next].arrival <= since(t0)) queue.push_back(next++);
    for (int32_t e = 0; e < nengines && !queue.empty(); ++e) {
      if (busy[(size_t)e]) continue;
      if ((rc = spws_batch_reset(engines[e]))) return rc;
      int32_t count = 0;
      int64_t prod = 0;
      while (!queue.empty() && count < cap_req) {
        const int32_t r = queue.front();
        const spws_request& q = reqs[r];
        if (count > 0 && prod + q.products > cap_prod) break;  // keep the batch short
        int32_t slot;
        rc = spws_batch_add(engines[e], q.m, q.k, q.n, q.A_pos, q.A_crd, q.A_vals, q.annz, q.B_pos, q.B_crd,
                            q.B_vals, q.bnnz, q.products, q.max_row_products, &slot);
        if (rc == SPWS_OK) {
          queue.pop_front();
          owned[(size_t)e].push_back(r);
          if (batch_of) batch_of[r] = (int32_t)batch_id;
          ++count;
          prod += q.products;
          continue;
        }
        if (rc == SPWS_ERR_OOM && count > 0) break;  // this batch is full
        if (rc == SPWS_ERR_OOM || rc == SPWS_ERR_UNSUPPORTED) {
          // no empty engine can hold it: run it alone through the plain path
          queue.pop_front();
          if ((rc = run_plain(ctx, q))) return rc;
          finish[r] = since(t0);
          if (batch_of) batch_of[r] = -1;
          ++done;
          ++st.fallback_requests;
          progress = true;
          continue;
        }
        return rc;
      }
      if (count > 0) {
        if ((rc = spws_batch_launch(engines[e]))) return rc;
        busy[(size_t)e] = true;
        ++batch_id;
        ++st.batches;
        progress = true;
      }
    }
    if (!progress) {
      bool any_busy = false;
      for (int32_t e = 0; e < nengines; ++e) any_busy = any_busy || busy[(size_t)e];
      if (!any_busy && queue.empty() && next < nreq) {
        wait_until(t0, reqs[next].arrival);
      } else {
        std::this_thread::yield();
      }
    }
  }
  st.wall_seconds = since(t0);
  if (stats) *stats = st;
  return SPWS_OK;
}

}  // extern "C"
