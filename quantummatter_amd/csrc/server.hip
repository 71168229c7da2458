// Request server of libqmhip.so: admission by HBM footprint, resident states,
// continuous batching of small widths, concurrent full-chip lanes.
//
// Requests (a circuit of 1..34 qubits plus what to return: the amplitudes,
// terminal shots from caller uniforms, or Pauli expectation values) are
// queued in arrival order.  A request runs once its state fits the memory
// budget beside everything already running.  Lanes are worker threads, each
// with its own context and stream, so their work overlaps on the device:
//   - the batch lane takes every waiting request of width <= batch_max_qubits
//     that fits and runs them in one launch of the batched kernel (one
//     workgroup per state, states resident for the launch), grouped by tile
//     shape; requests arriving meanwhile form the next launch;
//   - each full-chip lane takes the oldest wider request and runs it with the
//     fused tile passes on its own resident state; an idle full-chip lane also
//     takes a waiting small request while the batch lane is busy.
// Every request's result is computed by the same kernels as the single-state
// and batched entry points, so it is the same alone and in the mix (within
// the stated tolerances: amplitudes 1e-12 in c128, identical samples).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "qm_internal.hpp"

namespace qm {
namespace {

struct Req {
  int64_t id = 0;
  int qubits = 0, precision = QM_C128, kind = 0;
  std::vector<DevGate> gates;
  std::vector<double> uniforms;
  std::vector<qm_pauli> terms;
  void *out = nullptr;
  int64_t bytes = 0;  // device footprint (the state)
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = QM_OK;
  std::string msg;
  int lane = -1;
  int64_t fallbacks = 0;
  bool done = false;
};

struct Server {
  int device = 0;
  int64_t budget = 0;
  int batch_max_qubits = 16;
  int schedule = QM_SCHED_FUSED;  // full-chip lanes
  int max_batch = 1024;
  std::vector<Ctx *> lanes;  // [0] batch lane, [1..] full-chip lanes
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> small_q, large_q;
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
  int64_t next_id = 1;
  int64_t reserved = 0, peak = 0;
  int running = 0;
  bool batch_busy = false;
  bool stopping = false;
  std::vector<std::thread> threads;
  std::chrono::steady_clock::time_point t0;
  std::vector<double> latencies;
  int64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched = 0;
  double first_submit = -1, last_done = 0;

  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

int64_t footprint(int qubits, int precision) { return amp_bytes(precision) << qubits; }

// Finishes a request (lock held by the caller).
void finish(Server *sv, Req *r, int status, const std::string &msg) {
  r->status = status;
  r->msg = msg;
  r->t_done = sv->now();
  r->done = true;
  sv->completed++;
  if (status) sv->failed++;
  sv->latencies.push_back(r->t_done - r->t_submit);
  sv->last_done = std::max(sv->last_done, r->t_done);
}

// ExpectationValue for every term from FP64 numerators and the mass
// (StateVector.Observations.cs:107-141).
void expectation_values(const double *num, double mass, int64_t n, double *values) {
  if (!std::isfinite(mass) || !(mass > 0))
    fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
  for (int64_t i = 0; i < n; i++) {
    if (!std::isfinite(num[i])) fail(QM_ERR_NONFINITE, "The observable accumulation is nonfinite.");
    values[i] = num[i] / mass;
  }
}

// One request on a full-chip lane: resident state, fused passes, result.
void run_single(Ctx *c, Req *r, int schedule) {
  State *s = state_new(c, r->qubits, r->precision);
  try {
    launch_set_basis(s, 0, c->stream);
    run_gates(s, r->gates, schedule, nullptr);
    if (r->kind == QM_REQ_AMPLITUDES) {
      copy_to_host(c, r->out, s->data, (size_t)s->bytes);
    } else if (r->kind == QM_REQ_SHOTS) {
      r->fallbacks = cdf_search(s, r->uniforms.data(), (int64_t)r->uniforms.size(),
                                static_cast<int64_t *>(r->out), c->stream);
    } else {
      const double mass = reduce_mass(s, c->stream).total;
      if (!std::isfinite(mass) || !(mass > 0))
        fail(QM_ERR_MASS, "The state has no finite positive probability mass.");
      std::vector<double> num(r->terms.size());
      reduce_pauli(s, r->terms.data(), (int64_t)r->terms.size(), num.data(), c->stream);
      expectation_values(num.data(), mass, (int64_t)num.size(), static_cast<double *>(r->out));
    }
  } catch (...) {
    state_free(s);
    throw;
  }
  state_free(s);
}

// Same-shape small requests in one batched launch; statuses per request.
void run_batched(Ctx *c, const std::vector<Req *> &group, const FusedShape &shape) {
  BatchLaunch L;
  L.precision = group.front()->precision;
  L.shape = shape;
  L.keep = true;
  std::vector<double> shot_u;
  std::vector<TermGroups> tg(group.size());
  for (size_t i = 0; i < group.size(); i++) {
    Req *r = group[i];
    std::vector<BatchStep> steps(r->gates.size());
    for (size_t k = 0; k < r->gates.size(); k++) steps[k].gate = r->gates[k];
    BatchProgram p = plan_batch(steps, r->qubits, shape);
    BatchItem it{};
    it.state_off = L.keep_len;
    it.len = 1ull << r->qubits;
    it.tiles = 1ull << (r->qubits - shape.K);
    it.seg_begin = (int32_t)L.prog.segs.size();
    it.seg_count = (int32_t)p.segs.size();
    const int32_t pass_base = (int32_t)L.prog.passes.size();
    const int32_t op_base = (int32_t)L.prog.ops.size();
    for (BatchSeg sg : p.segs) {
      sg.pass_begin += pass_base;
      L.prog.segs.push_back(sg);
    }
    for (DevPass d : p.passes) {
      d.op_offset += op_base;
      L.prog.passes.push_back(d);
    }
    L.prog.ops.insert(L.prog.ops.end(), p.ops.begin(), p.ops.end());
    if (r->kind == QM_REQ_EXPECTATION) {
      tg[i] = group_terms(r->terms.data(), (int64_t)r->terms.size());
      it.group_begin = (int32_t)L.groups.size();
      it.group_count = (int32_t)tg[i].groups.size();
      it.numer_off = L.n_numer;
      it.want_mass = 1;
      L.groups.insert(L.groups.end(), tg[i].groups.begin(), tg[i].groups.end());
      L.n_numer += (int64_t)r->terms.size();
    }
    if (r->kind == QM_REQ_SHOTS) {
      it.shots = (int64_t)r->uniforms.size();
      it.shot_off = L.n_shots;
      shot_u.insert(shot_u.end(), r->uniforms.begin(), r->uniforms.end());
      L.n_shots += it.shots;
    }
    L.items.push_back(it);
    L.keep_len += (int64_t)it.len;
  }
  L.shot_u = shot_u.data();
  const int64_t ab = amp_bytes(L.precision);
  L.on_states = [&](const void *states, const double *, hipStream_t) {
    for (size_t i = 0; i < group.size(); i++)
      if (group[i]->kind == QM_REQ_AMPLITUDES)
        copy_to_host(c, group[i]->out, static_cast<const char *>(states) + L.items[i].state_off * ab,
                 (size_t)(L.items[i].len * ab));
  };
  run_batch_launch(c, L);
  for (size_t i = 0; i < group.size(); i++) {
    Req *r = group[i];
    const BatchItem &it = L.items[i];
    r->fallbacks = L.fallbacks;  // per launch (the kernel counts per launch)
    const int code = L.err[i];
    if (code) {
      r->status = code;
      r->msg = code == QM_ERR_MASS ? "The state has no finite positive probability mass."
                                   : "batched request failed";
      continue;
    }
    try {
      if (r->kind == QM_REQ_SHOTS) {
        std::memcpy(r->out, L.samples.data() + it.shot_off, (size_t)it.shots * 8);
      } else if (r->kind == QM_REQ_EXPECTATION) {
        const int64_t nt = (int64_t)r->terms.size();
        std::vector<double> num((size_t)nt);
        for (int64_t k = 0; k < nt; k++)
          num[(size_t)tg[i].order[(size_t)k]] = L.numer[(size_t)(it.numer_off + k)];
        expectation_values(num.data(), L.mass[i], nt, static_cast<double *>(r->out));
      }
    } catch (const Status &st) {
      r->status = st.code;
      r->msg = st.msg;
    }
  }
}

void lane_full(Server *sv, int lane) {
  (void)hipSetDevice(sv->device);
  Ctx *c = sv->lanes[(size_t)lane];
  for (;;) {
    Req *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] {
        if (sv->stopping) return true;
        if (!sv->large_q.empty() && sv->reserved + sv->large_q.front()->bytes <= sv->budget) return true;
        if (sv->large_q.empty() && sv->batch_busy && !sv->small_q.empty() &&
            sv->reserved + sv->small_q.front()->bytes <= sv->budget)
          return true;
        // A request that cannot fit even on an idle device fails instead of waiting forever.
        if (!sv->large_q.empty() && sv->running == 0 && sv->large_q.front()->bytes > sv->budget) return true;
        return false;
      });
      if (sv->stopping && sv->large_q.empty()) return;
      if (!sv->large_q.empty() && sv->large_q.front()->bytes > sv->budget && sv->running == 0) {
        Req *bad = sv->large_q.front();
        sv->large_q.pop_front();
        finish(sv, bad, QM_ERR_BUDGET, "Request exceeds the device memory budget.");
        sv->cv_done.notify_all();
        continue;
      }
      if (!sv->large_q.empty() && sv->reserved + sv->large_q.front()->bytes <= sv->budget) {
        r = sv->large_q.front();
        sv->large_q.pop_front();
      } else if (!sv->small_q.empty() && sv->reserved + sv->small_q.front()->bytes <= sv->budget) {
        r = sv->small_q.front();
        sv->small_q.pop_front();
      } else {
        continue;
      }
      sv->reserved += r->bytes;
      sv->peak = std::max(sv->peak, sv->reserved);
      sv->running++;
      r->t_start = sv->now();
      r->lane = lane;
    }
    int code = QM_OK;
    std::string msg;
    try {
      run_single(c, r, sv->schedule);
    } catch (const Status &s) {
      code = s.code;
      msg = s.msg;
    } catch (const std::exception &e) {
      code = QM_ERR_DEVICE;
      msg = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->reserved -= r->bytes;
      sv->running--;
      finish(sv, r, code, msg);
    }
    sv->cv_done.notify_all();
    sv->cv_work.notify_all();
  }
}

void lane_batch(Server *sv) {
  (void)hipSetDevice(sv->device);
  Ctx *c = sv->lanes[0];
  for (;;) {
    std::vector<Req *> take;
    int64_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] {
        return sv->stopping || (!sv->small_q.empty() &&
                                sv->reserved + sv->small_q.front()->bytes <= sv->budget);
      });
      if (sv->stopping && sv->small_q.empty()) return;
      while (!sv->small_q.empty() && (int)take.size() < sv->max_batch &&
             sv->reserved + bytes + sv->small_q.front()->bytes <= sv->budget) {
        bytes += sv->small_q.front()->bytes;
        take.push_back(sv->small_q.front());
        sv->small_q.pop_front();
      }
      if (take.empty()) continue;
      sv->reserved += bytes;
      sv->peak = std::max(sv->peak, sv->reserved);
      sv->running++;
      sv->batch_busy = true;
      const double t = sv->now();
      for (Req *r : take) {
        r->t_start = t;
        r->lane = 0;
      }
    }
    // Widths below the batched kernels' minimum and lone requests run on the
    // full-chip path (lower latency for one state); the rest by tile shape.
    std::vector<Req *> single;
    std::vector<std::pair<FusedShape, std::vector<Req *>>> groups;
    for (Req *r : take) {
      const FusedShape sh = batch_shape(r->precision, r->qubits);
      if (sh.K == 0 || take.size() == 1) {
        single.push_back(r);
        continue;
      }
      bool placed = false;
      for (auto &g : groups)
        if (g.first.K == sh.K && g.first.L == sh.L && g.second.front()->precision == r->precision) {
          g.second.push_back(r);
          placed = true;
          break;
        }
      if (!placed) groups.push_back({sh, {r}});
    }
    for (Req *r : single) {
      try {
        run_single(c, r, sv->schedule);
      } catch (const Status &s) {
        r->status = s.code;
        r->msg = s.msg;
      } catch (const std::exception &e) {
        r->status = QM_ERR_DEVICE;
        r->msg = e.what();
      }
    }
    int64_t launched = 0;
    for (auto &g : groups) {
      try {
        run_batched(c, g.second, g.first);
        launched++;
      } catch (const Status &s) {
        for (Req *r : g.second) {
          r->status = s.code;
          r->msg = s.msg;
        }
      } catch (const std::exception &e) {
        for (Req *r : g.second) {
          r->status = QM_ERR_DEVICE;
          r->msg = e.what();
        }
      }
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->reserved -= bytes;
      sv->running--;
      sv->batch_busy = false;
      sv->batches += launched;
      for (auto &g : groups) sv->batched += (int64_t)g.second.size();
      for (Req *r : take) finish(sv, r, r->status, r->msg);
    }
    sv->cv_done.notify_all();
    sv->cv_work.notify_all();
  }
}

Server *as_server(qm_server *h) {
  if (!h) fail(QM_ERR_ARG, "server is null");
  return reinterpret_cast<Server *>(h);
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

int qm_server_create(const qm_server_options *opt, qm_server **out) {
  return guarded([&] {
    if (!out) fail(QM_ERR_ARG, "out is null");
    *out = nullptr;
    qm_server_options o{};
    if (opt) o = *opt;
    if (o.full_lanes < 0 || o.full_lanes > 8) fail(QM_ERR_RANGE, "full_lanes must be 0..8");
    if (o.max_batch < 0) fail(QM_ERR_RANGE, "max_batch must be nonnegative");
    auto sv = std::make_unique<Server>();
    sv->device = o.device_index;
    if (o.schedule != 0 && o.schedule != QM_SCHED_FUSED && o.schedule != QM_SCHED_AUTO)
      fail(QM_ERR_ARG, "schedule must be 0, QM_SCHED_FUSED or QM_SCHED_AUTO");
    sv->schedule = o.schedule == QM_SCHED_AUTO ? QM_SCHED_AUTO : QM_SCHED_FUSED;
    sv->batch_max_qubits = o.batch_max_qubits > 0 ? o.batch_max_qubits : 14;
    sv->max_batch = o.max_batch > 0 ? o.max_batch : 1024;
    const int full = o.full_lanes > 0 ? o.full_lanes : 3;
    try {
      for (int i = 0; i < 1 + full; i++) {
        qm_ctx_options co{};
        co.device_index = o.device_index;
        qm_ctx *h = nullptr;
        const int rc = qm_ctx_create(&co, &h);
        if (rc) fail(rc, qm_last_error());
        sv->lanes.push_back(reinterpret_cast<Ctx *>(h));
      }
      const int64_t mem = sv->lanes.front()->device_memory;
      // Default budget: the device memory less a reserve for scratch and programs.
      sv->budget = o.device_budget_bytes > 0 ? std::min<int64_t>(o.device_budget_bytes, mem)
                                             : mem - ((int64_t)4 << 30);
      sv->t0 = std::chrono::steady_clock::now();
      Server *raw = sv.get();
      raw->threads.emplace_back(lane_batch, raw);
      for (int i = 1; i <= full; i++) raw->threads.emplace_back(lane_full, raw, i);
    } catch (...) {
      for (Ctx *c : sv->lanes) qm_ctx_destroy(reinterpret_cast<qm_ctx *>(c));
      throw;
    }
    *out = reinterpret_cast<qm_server *>(sv.release());
  });
}

int qm_server_submit(qm_server *h, const qm_request *rq, int64_t *id) {
  return guarded([&] {
    Server *sv = as_server(h);
    if (!rq || !id) fail(QM_ERR_ARG, "request is null");
    if (rq->qubits < 1 || rq->qubits > 34) fail(QM_ERR_RANGE, "qubits must be between 1 and 34.");
    if (rq->precision != QM_C128 && rq->precision != QM_C64) fail(QM_ERR_ARG, "unknown precision");
    if (rq->kind < QM_REQ_AMPLITUDES || rq->kind > QM_REQ_EXPECTATION) fail(QM_ERR_ARG, "unknown request kind");
    if (!rq->out) fail(QM_ERR_ARG, "output buffer is null");
    auto r = std::make_unique<Req>();
    r->qubits = rq->qubits;
    r->precision = rq->precision;
    r->kind = rq->kind;
    r->gates = prepare_all(rq->ops, rq->n_ops, rq->qubits);
    if (rq->kind == QM_REQ_SHOTS) {
      if (rq->shots < 1 || !rq->uniforms) fail(QM_ERR_RANGE, "shots must be positive");
      r->uniforms.assign(rq->uniforms, rq->uniforms + rq->shots);
    }
    if (rq->kind == QM_REQ_EXPECTATION) {
      if (rq->n_terms < 0 || (rq->n_terms > 0 && !rq->terms)) fail(QM_ERR_ARG, "terms are null");
      const uint64_t all = rq->qubits >= 64 ? ~0ull : ((1ull << rq->qubits) - 1);
      for (int64_t i = 0; i < rq->n_terms; i++)
        if ((rq->terms[i].flip_mask | rq->terms[i].sign_mask) & ~all)
          fail(QM_ERR_ARG, "Observable and state sizes must match.");
      r->terms.assign(rq->terms, rq->terms + rq->n_terms);
    }
    r->out = rq->out;
    r->bytes = footprint(rq->qubits, rq->precision);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stopping) fail(QM_ERR_DISPOSED, "The server has been stopped.");
    r->id = sv->next_id++;
    r->t_submit = sv->now();
    if (sv->first_submit < 0) sv->first_submit = r->t_submit;
    sv->submitted++;
    Req *raw = r.get();
    sv->reqs[raw->id] = std::move(r);
    *id = raw->id;
    if (raw->bytes > sv->budget) {
      // Can never be admitted, whichever lane would take it: fails now (its
      // wait reports the budget error) instead of waiting in a queue forever.
      finish(sv, raw, QM_ERR_BUDGET, "Request exceeds the device memory budget.");
      sv->cv_done.notify_all();
      return;
    }
    if (raw->qubits <= sv->batch_max_qubits)
      sv->small_q.push_back(raw);
    else
      sv->large_q.push_back(raw);
    sv->cv_work.notify_all();
  });
}

int qm_server_wait(qm_server *h, int64_t id, qm_request_status *status) {
  int code = QM_OK;
  std::string msg;
  const int rc = guarded([&] {
    Server *sv = as_server(h);
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) fail(QM_ERR_ARG, "unknown request id");
    Req *r = it->second.get();
    sv->cv_done.wait(lk, [&] { return r->done; });
    if (status) {
      status->id = r->id;
      status->status = r->status;
      status->lane = r->lane;
      status->submit_ms = r->t_submit;
      status->start_ms = r->t_start;
      status->done_ms = r->t_done;
      status->fallbacks = r->fallbacks;
    }
    code = r->status;
    msg = r->msg;
    sv->reqs.erase(it);
  });
  if (rc) return rc;
  if (code) set_error(msg);
  return code;
}

int qm_server_drain(qm_server *h) {
  return guarded([&] {
    Server *sv = as_server(h);
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->completed == sv->submitted; });
  });
}

int qm_server_get_stats(qm_server *h, qm_server_stats *out) {
  return guarded([&] {
    Server *sv = as_server(h);
    if (!out) fail(QM_ERR_ARG, "out is null");
    std::lock_guard<std::mutex> lk(sv->mu);
    qm_server_stats s{};
    s.submitted = sv->submitted;
    s.completed = sv->completed;
    s.failed = sv->failed;
    s.pending = (int64_t)(sv->small_q.size() + sv->large_q.size());
    s.reserved_bytes = sv->reserved;
    s.peak_reserved_bytes = sv->peak;
    s.budget_bytes = sv->budget;
    s.batches = sv->batches;
    s.batched_requests = sv->batched;
    s.p50_ms = percentile(sv->latencies, 0.50);
    s.p99_ms = percentile(sv->latencies, 0.99);
    s.span_ms = sv->first_submit < 0 ? 0 : sv->last_done - sv->first_submit;
    *out = s;
  });
}

int qm_server_reset_stats(qm_server *h) {
  return guarded([&] {
    Server *sv = as_server(h);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->latencies.clear();
    sv->submitted -= sv->completed;
    sv->completed = 0;
    sv->failed = 0;
    sv->batches = 0;
    sv->batched = 0;
    sv->peak = sv->reserved;
    sv->first_submit = -1;
    sv->last_done = 0;
  });
}

int qm_server_destroy(qm_server *h) {
  return guarded([&] {
    if (!h) return;
    Server *sv = reinterpret_cast<Server *>(h);
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_done.wait(lk, [&] { return sv->completed == sv->submitted; });
      sv->stopping = true;
    }
    sv->cv_work.notify_all();
    for (std::thread &t : sv->threads) t.join();
    for (Ctx *c : sv->lanes) qm_ctx_destroy(reinterpret_cast<qm_ctx *>(c));
    delete sv;
  });
}

}  // extern "C"
