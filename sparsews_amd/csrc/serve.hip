// Request loop for SpGEMM serving: a plain mode (one row-wise call per
// request, in arrival order) and a continuously batched mode over a set of
// batch engines.
//
// Batched mode, per iteration of the loop:
//   1. completions: every engine whose batch finished releases its requests
//      (their finish time is the time the loop observes the completion, after
//      the batch has copied each result into its request's output buffer);
//   2. arrivals: requests whose arrival time has passed join a queue: the
//      large queue when their upper bound of output entries is at least
//      large_products, the small queue otherwise;
//   3. admission: each idle engine of the small lane takes queued small
//      requests in arrival order while its pools hold them (rows, operand
//      nonzeros, and the upper bound of the output entries) and the batch
//      stays under the products cap, then launches its graph.  Each idle
//      engine of the large lane takes one large request.  An idle small-lane
//      engine with no small request waiting takes a large one, so no engine
//      idles while work waits.  Small requests therefore never queue behind
//      a large one, and requests admitted while other engines run overlap
//      with them on the device (one stream per engine).  A request no
//      engine can ever hold runs through the plain path.
#include <chrono>
#include <deque>
#include <thread>
#include <utility>
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

// One row-wise call; the result is copied into the request's output buffer
// when it has one.  *nnz_out = the output nnz (-1: it exceeds C_cap).
int run_plain(spws_ctx* ctx, const spws_request& q, int64_t* nnz_out) {
  int64_t nnz;
  int32_t *p, *c;
  double* v;
  int rc = spws_spgemm_rowwise(ctx, q.m, q.k, q.n, q.A_pos, q.A_crd, q.A_vals, q.B_pos, q.B_crd, q.B_vals, &nnz, &p,
                               &c, &v);
  if (rc) return rc;
  *nnz_out = nnz;
  if (!q.C_pos) return SPWS_OK;
  if (nnz > q.C_cap) {
    *nnz_out = -1;
    return SPWS_OK;
  }
  SPWS_HIP(hipMemcpyAsync(q.C_pos, p, ((size_t)q.m + 1) * 4, hipMemcpyDeviceToDevice, ctx->stream));
  if (nnz > 0) {
    SPWS_HIP(hipMemcpyAsync(q.C_crd, c, (size_t)nnz * 4, hipMemcpyDeviceToDevice, ctx->stream));
    SPWS_HIP(hipMemcpyAsync(q.C_vals, v, (size_t)nnz * 8, hipMemcpyDeviceToDevice, ctx->stream));
  }
  SPWS_HIP(hipStreamSynchronize(ctx->stream));
  return SPWS_OK;
}

}  // namespace

extern "C" {

int spws_serve(spws_ctx* ctx, spws_batch** engines, int32_t nengines, int32_t nreq, const spws_request* reqs,
               const spws_serve_options* opt, double* finish, int32_t* batch_of, int64_t* result_nnz,
               spws_serve_stats* stats) {
  SPWS_REQUIRE(ctx && reqs && opt && finish && nreq >= 0, SPWS_ERR_INVALID, "null argument");
  for (int32_t i = 1; i < nreq; ++i)
    SPWS_REQUIRE(reqs[i].arrival >= reqs[i - 1].arrival, SPWS_ERR_INVALID, "requests must be sorted by arrival");
  for (int32_t i = 0; i < nreq; ++i)
    SPWS_REQUIRE(!reqs[i].C_pos || reqs[i].C_cap == 0 || (reqs[i].C_crd && reqs[i].C_vals), SPWS_ERR_INVALID,
                 "request %d: an output buffer with capacity needs columns and values", i);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  spws_serve_stats st{};
  int64_t nnz_tmp = 0;
  auto nnz_slot = [&](int32_t r) -> int64_t* { return result_nnz ? &result_nnz[r] : &nnz_tmp; };
  const Clock::time_point t0 = Clock::now();
  if (opt->mode == 0) {
    for (int32_t r = 0; r < nreq; ++r) {
      wait_until(t0, reqs[r].arrival);
      if ((rc = run_plain(ctx, reqs[r], nnz_slot(r)))) return rc;
      finish[r] = since(t0);
      if (batch_of) batch_of[r] = -1;
    }
    st.wall_seconds = since(t0);
    if (stats) *stats = st;
    return SPWS_OK;
  }
  SPWS_REQUIRE(engines && nengines >= 1, SPWS_ERR_INVALID, "batched mode needs at least one engine");
  SPWS_REQUIRE(opt->large_engines >= 0 && opt->large_engines < nengines, SPWS_ERR_INVALID,
               "large_engines must leave at least one engine for small requests");
  const int32_t cap_req = opt->max_batch_requests > 0 ? opt->max_batch_requests : INT32_MAX;
  const int64_t cap_prod = opt->max_batch_products > 0 ? opt->max_batch_products : INT64_MAX;
  const int64_t large_at = opt->large_products > 0 ? opt->large_products : INT64_MAX;
  // engines [0, n_small) form the small lane, the rest the large lane
  const int32_t n_small = nengines - (opt->large_products > 0 ? opt->large_engines : 0);
  std::vector<std::vector<std::pair<int32_t, int32_t>>> owned((size_t)nengines);  // (request, slot)
  std::vector<bool> busy((size_t)nengines, false);
  std::deque<int32_t> small_q, large_q;
  int32_t next = 0, done = 0;
  int64_t batch_id = 0;

  auto plain_fallback = [&](int32_t r) -> int {
    const int rc2 = run_plain(ctx, reqs[r], nnz_slot(r));
    if (rc2) return rc2;
    finish[r] = since(t0);
    if (batch_of) batch_of[r] = -1;
    ++done;
    ++st.fallback_requests;
    return SPWS_OK;
  };
  // fills engine e from queue q (at most `limit` requests) and launches it;
  // *launched says whether a batch started
  auto fill = [&](int32_t e, std::deque<int32_t>& q, int32_t limit, bool* launched) -> int {
    *launched = false;
    int rc2;
    if ((rc2 = spws_batch_reset(engines[e]))) return rc2;
    int32_t count = 0;
    int64_t prod = 0;
    while (!q.empty() && count < limit) {
      const int32_t r = q.front();
      const spws_request& x = reqs[r];
      if (count > 0 && prod + x.products > cap_prod) break;  // keep the batch short
      int32_t slot;
      rc2 = spws_batch_add(engines[e], x.m, x.k, x.n, x.A_pos, x.A_crd, x.A_vals, x.annz, x.B_pos, x.B_crd,
                           x.B_vals, x.bnnz, x.products, x.max_row_products, &slot);
      if (rc2 == SPWS_OK) {
        if (x.C_pos && (rc2 = spws_batch_set_output(engines[e], slot, x.C_pos, x.C_crd, x.C_vals, x.C_cap)))
          return rc2;
        q.pop_front();
        owned[(size_t)e].push_back({r, slot});
        if (batch_of) batch_of[r] = (int32_t)batch_id;
        ++count;
        prod += x.products;
        continue;
      }
      if (rc2 == SPWS_ERR_OOM && count > 0) break;  // this batch is full
      if (rc2 == SPWS_ERR_OOM || rc2 == SPWS_ERR_UNSUPPORTED) {
        // an empty engine cannot hold it: run it alone through the plain path
        q.pop_front();
        if ((rc2 = plain_fallback(r))) return rc2;
        continue;
      }
      return rc2;
    }
    if (count > 0) {
      if ((rc2 = spws_batch_launch(engines[e]))) return rc2;
      busy[(size_t)e] = true;
      ++batch_id;
      ++st.batches;
      *launched = true;
    }
    return SPWS_OK;
  };

  while (done < nreq) {
    bool progress = false;
    for (int32_t e = 0; e < nengines; ++e) {
      if (!busy[(size_t)e]) continue;
      int fin = 0;
      if ((rc = spws_batch_poll(engines[e], &fin))) return rc;
      if (fin) {
        const double t = since(t0);
        for (const auto& rs : owned[(size_t)e]) {
          finish[rs.first] = t;
          if ((rc = spws_batch_result_nnz(engines[e], rs.second, nnz_slot(rs.first)))) return rc;
        }
        done += (int32_t)owned[(size_t)e].size();
        owned[(size_t)e].clear();
        busy[(size_t)e] = false;
        progress = true;
      }
    }
    while (next < nreq && reqs[next].arrival <= since(t0)) {
      (reqs[next].products >= large_at ? large_q : small_q).push_back(next);
      ++next;
    }
    for (int32_t e = n_small; e < nengines && !large_q.empty(); ++e) {
      if (busy[(size_t)e]) continue;
      bool launched;
      if ((rc = fill(e, large_q, 1, &launched))) return rc;
      progress = progress || launched;
    }
    for (int32_t e = 0; e < n_small && (!small_q.empty() || !large_q.empty()); ++e) {
      if (busy[(size_t)e]) continue;
      bool launched;
      if (!small_q.empty()) {
        if ((rc = fill(e, small_q, cap_req, &launched))) return rc;
      } else {
        if ((rc = fill(e, large_q, 1, &launched))) return rc;
      }
      progress = progress || launched;
    }
    if (!progress) {
      bool any_busy = false;
      for (int32_t e = 0; e < nengines; ++e) any_busy = any_busy || busy[(size_t)e];
      if (!any_busy && small_q.empty() && large_q.empty() && next < nreq) {
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
