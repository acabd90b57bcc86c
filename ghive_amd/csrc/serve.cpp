// Serving form: a request loop in front of the operators. Requests (fragments, group-bys, sorts)
// arrive from any host thread -- in GHive, from the JNI bridge of concurrent Tez tasks -- and are
// admitted in arrival order as soon as a worker slot is free AND their HBM footprint (inputs plus
// intermediates) fits the budget next to the requests already in flight. Every slot owns a
// context (stream + memory manager) on the same device, so admitted requests run concurrently
// and overlap on the GPU; dimension hash tables registered with the server stay resident across
// requests. No counterpart in the reference, which executes one task per JNI call.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ghive/ghive.h"

namespace {

double now_us() {
  using namespace std::chrono;
  return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

uint64_t table_bytes(const ghive_table *t) {
  uint64_t b = 0;
  if (!t) return 0;
  for (uint32_t c = 0; c < t->ncols; ++c) {
    int ty = t->columns[c].type;
    b += t->row_num * ((ty == GHIVE_LONG || ty == GHIVE_DOUBLE) ? 8 : 4);
  }
  return b;
}

struct Job {
  uint64_t id;
  ghive_request req;
  uint64_t footprint;
  double submit_us;
};

}  // namespace

struct ghive_server {
  int device;
  ghive_server_config cfg;
  std::vector<ghive_ctx *> slots;
  std::vector<std::thread> workers;
  std::thread dispatcher;
  std::mutex mu;
  std::condition_variable cv_queue;   // dispatcher: new job or capacity freed
  std::condition_variable cv_slot;    // workers: job assigned
  std::condition_variable cv_done;    // waiters: a job finished
  std::deque<Job> queue;
  std::vector<int> slot_busy;
  std::vector<Job> slot_job;
  std::map<uint64_t, ghive_response> done;
  std::set<uint64_t> pending;  // submitted and not yet waited for
  uint64_t next_id = 1;
  uint64_t inflight_bytes = 0;
  uint32_t inflight = 0;
  bool stop = false;
  // statistics
  uint64_t completed = 0, rejected = 0, peak_inflight_bytes = 0;
  uint32_t peak_inflight = 0;
};

namespace {

uint64_t type_bytes(int t) { return (t == GHIVE_LONG || t == GHIVE_DOUBLE) ? 8 : 4; }

// Bytes of a group table holding `capacity` groups with `naggs` aggregates (slots = next power of
// two >= 2 * capacity; one key word and at most five accumulator words per aggregate).
uint64_t group_table_bytes(uint64_t capacity, uint32_t naggs) {
  uint64_t slots = 64;
  while (slots < capacity * 2) slots <<= 1;
  return (slots + 1) * 8 * (1 + 5ull * naggs);
}

uint64_t out_bytes(const ghive_request &r) {
  switch (r.kind) {
    case GHIVE_REQ_FRAGMENT:
      return r.fragment ? r.capacity * (4ull * r.fragment->nkeys + 8ull * r.fragment->naggs) : 0;
    case GHIVE_REQ_GROUPBY: {
      uint64_t b = 0;
      for (uint32_t k = 0; k < r.nkeys && r.table; ++k) b += type_bytes(r.table->columns[r.keys[k]].type);
      return r.capacity * (b + 8ull * r.naggs);
    }
    case GHIVE_REQ_SORT: return r.table ? 4ull * r.table->row_num : 0;
    default: return 0;
  }
}

// HBM footprint of a request: what it allocates after admission (inputs and outputs too when
// they arrive in host memory) plus the intermediates of its operator.
uint64_t footprint_of(const ghive_request &r) {
  if (r.footprint_bytes) return r.footprint_bytes;
  const ghive_table *in = r.kind == GHIVE_REQ_FRAGMENT ? (r.fragment ? &r.fragment->fact : nullptr) : r.table;
  uint64_t io = r.host_io ? table_bytes(in) + out_bytes(r) : 0;
  uint64_t n = in ? in->row_num : 0;
  switch (r.kind) {
    case GHIVE_REQ_FRAGMENT:
      return io + group_table_bytes(r.capacity, r.fragment ? r.fragment->naggs : 0) + (1 << 20);
    case GHIVE_REQ_GROUPBY:
      return io + group_table_bytes(r.capacity, r.naggs) + (1 << 20);
    case GHIVE_REQ_SORT: {
      uint64_t kw = 0;
      for (uint32_t k = 0; k < r.nkeys && r.table; ++k) kw += type_bytes(r.table->columns[r.keys[k]].type);
      uint64_t words = (kw + 7) / 8;
      return io + n * (16 * words + 4) + (n / 1024 + 1) * 1024 + (1 << 20);
    }
    default: return 0;
  }
}

// Device copies of a request's host inputs and outputs, released when the request ends.
struct Staging {
  ghive_ctx *ctx;
  std::vector<void *> bufs;
  explicit Staging(ghive_ctx *c) : ctx(c) {}
  ~Staging() {
    for (void *p : bufs) ghive_free(ctx, p);
  }
  int alloc(uint64_t bytes, void **out) {
    int s = ghive_malloc(ctx, bytes ? bytes : 1, out);
    if (s == GHIVE_OK) bufs.push_back(*out);
    return s;
  }
  // Uploads every column of a host table into `cols` and points `dev` at the copies.
  int upload(const ghive_table *host, std::vector<ghive_column> &cols, ghive_table *dev) {
    *dev = *host;
    cols.assign(host->columns, host->columns + host->ncols);
    for (auto &c : cols) {
      uint64_t bytes = host->row_num * type_bytes(c.type);
      void *d = nullptr;
      int s = alloc(bytes, &d);
      if (s != GHIVE_OK) return s;
      if (bytes) {
        s = ghive_memcpy(ctx, d, c.data, bytes, 0);
        if (s != GHIVE_OK) return s;
      }
      c.data = d;
    }
    dev->columns = cols.data();
    return GHIVE_OK;
  }
  int outputs(const ghive_column *host, uint32_t ncols, uint64_t rows, std::vector<ghive_column> &cols) {
    cols.assign(host, host + ncols);
    for (auto &c : cols) {
      void *d = nullptr;
      int s = alloc(rows * type_bytes(c.type), &d);
      if (s != GHIVE_OK) return s;
      c.data = d;
    }
    return GHIVE_OK;
  }
  int download(ghive_column *host, const std::vector<ghive_column> &cols, uint64_t rows) {
    for (size_t i = 0; i < cols.size(); ++i) {
      uint64_t bytes = rows * type_bytes(cols[i].type);
      if (!bytes) continue;
      int s = ghive_memcpy(ctx, host[i].data, cols[i].data, bytes, 1);
      if (s != GHIVE_OK) return s;
    }
    return GHIVE_OK;
  }
};

int run_request(ghive_ctx *ctx, const ghive_request &r, uint64_t *groups) {
  *groups = 0;
  if (!r.host_io) {
    switch (r.kind) {
      case GHIVE_REQ_FRAGMENT:
        return ghive_fragment_execute(ctx, r.fragment, r.out_cols, r.capacity, groups);
      case GHIVE_REQ_GROUPBY:
        return ghive_groupby(ctx, r.table, nullptr, r.table ? r.table->row_num : 0, r.keys, r.nkeys, r.aggs,
                             r.naggs, r.out_cols, r.capacity, groups);
      case GHIVE_REQ_SORT: {
        int s = ghive_sort(ctx, r.table, r.keys, r.orders, r.nkeys, r.out_perm);
        if (s == GHIVE_OK) *groups = r.table->row_num;
        return s;
      }
      default: return GHIVE_ERR_INVALID;
    }
  }
  Staging st(ctx);
  std::vector<ghive_column> in_cols, outs;
  ghive_table dev_table;
  int s;
  switch (r.kind) {
    case GHIVE_REQ_FRAGMENT: {
      if (!r.fragment || (r.capacity && !r.out_cols)) return GHIVE_ERR_INVALID;
      ghive_fragment f = *r.fragment;
      if ((s = st.upload(&r.fragment->fact, in_cols, &f.fact)) != GHIVE_OK) return s;
      uint32_t nout = f.nkeys + f.naggs;
      if (r.capacity && (s = st.outputs(r.out_cols, nout, r.capacity, outs)) != GHIVE_OK) return s;
      s = ghive_fragment_execute(ctx, &f, r.capacity ? outs.data() : nullptr, r.capacity, groups);
      if (s != GHIVE_OK) return s;
      return st.download(r.out_cols, outs, *groups);
    }
    case GHIVE_REQ_GROUPBY: {
      if (!r.table || (r.capacity && !r.out_cols)) return GHIVE_ERR_INVALID;
      if ((s = st.upload(r.table, in_cols, &dev_table)) != GHIVE_OK) return s;
      if (r.capacity && (s = st.outputs(r.out_cols, r.nkeys + r.naggs, r.capacity, outs)) != GHIVE_OK) return s;
      s = ghive_groupby(ctx, &dev_table, nullptr, dev_table.row_num, r.keys, r.nkeys, r.aggs, r.naggs,
                        r.capacity ? outs.data() : nullptr, r.capacity, groups);
      if (s != GHIVE_OK) return s;
      return st.download(r.out_cols, outs, *groups);
    }
    case GHIVE_REQ_SORT: {
      if (!r.table || !r.out_perm) return GHIVE_ERR_INVALID;
      // only the key columns are needed on the device
      ghive_table keys_only = *r.table;
      std::vector<ghive_column> kc;
      std::vector<uint32_t> kidx(r.nkeys);
      for (uint32_t k = 0; k < r.nkeys; ++k) {
        if (r.keys[k] >= r.table->ncols) return GHIVE_ERR_INVALID;
        kc.push_back(r.table->columns[r.keys[k]]);
        kidx[k] = k;
      }
      keys_only.columns = kc.data();
      keys_only.ncols = r.nkeys;
      keys_only.key_num = 0;
      if ((s = st.upload(&keys_only, in_cols, &dev_table)) != GHIVE_OK) return s;
      void *perm = nullptr;
      uint64_t n = dev_table.row_num;
      if ((s = st.alloc(4 * n, &perm)) != GHIVE_OK) return s;
      s = ghive_sort(ctx, &dev_table, kidx.data(), r.orders, r.nkeys, static_cast<uint32_t *>(perm));
      if (s != GHIVE_OK) return s;
      if (n && (s = ghive_memcpy(ctx, r.out_perm, perm, 4 * n, 1)) != GHIVE_OK) return s;
      *groups = n;
      return GHIVE_OK;
    }
    default: return GHIVE_ERR_INVALID;
  }
}

void worker_main(ghive_server *S, uint32_t slot) {
  ghive_ctx *ctx = S->slots[slot];
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_slot.wait(lk, [&] { return S->stop || S->slot_busy[slot] == 1; });
      if (S->slot_busy[slot] != 1) return;  // stopping and nothing assigned
      job = S->slot_job[slot];
      S->slot_busy[slot] = 2;  // running
    }
    ghive_response resp;
    memset(&resp, 0, sizeof resp);
    resp.submit_us = job.submit_us;
    resp.start_us = now_us();
    resp.slot = slot;
    uint64_t groups = 0;
    resp.status = run_request(ctx, job.req, &groups);
    // every output is complete before the request is reported done
    if (resp.status == GHIVE_OK) resp.status = ghive_sync(ctx);
    resp.end_us = now_us();
    resp.groups = groups;
    if (resp.status != GHIVE_OK) snprintf(resp.error, sizeof resp.error, "%s", ghive_last_error(ctx));
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->done[job.id] = resp;
      S->inflight_bytes -= job.footprint;
      S->inflight -= 1;
      S->slot_busy[slot] = 0;
      S->completed += 1;
    }
    S->cv_queue.notify_all();
    S->cv_done.notify_all();
  }
}

void dispatcher_main(ghive_server *S) {
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_queue.wait(lk, [&] {
      if (S->stop) return true;
      if (S->queue.empty()) return false;
      const Job &head = S->queue.front();
      bool slot_free = false;
      for (int b : S->slot_busy) slot_free |= b == 0;
      return slot_free && S->inflight_bytes + head.footprint <= S->cfg.hbm_budget_bytes;
    });
    if (S->stop) return;
    // admit the head of the queue (arrival order) into the first free slot
    Job job = S->queue.front();
    S->queue.pop_front();
    uint32_t slot = 0;
    while (S->slot_busy[slot] != 0) ++slot;
    S->slot_job[slot] = job;
    S->slot_busy[slot] = 1;
    S->inflight_bytes += job.footprint;
    S->inflight += 1;
    if (S->inflight_bytes > S->peak_inflight_bytes) S->peak_inflight_bytes = S->inflight_bytes;
    if (S->inflight > S->peak_inflight) S->peak_inflight = S->inflight;
    S->cv_slot.notify_all();
  }
}

}  // namespace

extern "C" {

int ghive_server_create(int device, const ghive_server_config *cfg, ghive_server **out) {
  if (!cfg || !out) return GHIVE_ERR_INVALID;
  *out = nullptr;
  if (cfg->slots < 1 || cfg->slots > 16 || cfg->hbm_budget_bytes == 0) return GHIVE_ERR_INVALID;
  auto *S = new ghive_server();
  S->device = device;
  S->cfg = *cfg;
  S->slot_busy.assign(cfg->slots, 0);
  S->slot_job.resize(cfg->slots);
  for (uint32_t i = 0; i < cfg->slots; ++i) {
    ghive_ctx *c = nullptr;
    int st = ghive_init(device, cfg->slot_mem_limit_bytes, &c);
    if (st != GHIVE_OK) {
      for (ghive_ctx *x : S->slots) ghive_shutdown(x);
      delete S;
      return st;
    }
    S->slots.push_back(c);
  }
  for (uint32_t i = 0; i < cfg->slots; ++i) S->workers.emplace_back(worker_main, S, i);
  S->dispatcher = std::thread(dispatcher_main, S);
  *out = S;
  return GHIVE_OK;
}

int ghive_server_submit(ghive_server *S, const ghive_request *req, uint64_t *id) {
  if (!S || !req || !id) return GHIVE_ERR_INVALID;
  if (req->kind < GHIVE_REQ_FRAGMENT || req->kind > GHIVE_REQ_SORT) return GHIVE_ERR_INVALID;
  Job job;
  job.req = *req;
  job.footprint = footprint_of(*req);
  job.submit_us = now_us();
  std::lock_guard<std::mutex> lk(S->mu);
  if (job.footprint > S->cfg.hbm_budget_bytes) {
    S->rejected += 1;
    return GHIVE_ERR_ADMISSION;  // can never be admitted
  }
  job.id = S->next_id++;
  *id = job.id;
  S->pending.insert(job.id);
  S->queue.push_back(job);
  S->cv_queue.notify_all();
  return GHIVE_OK;
}

int ghive_server_wait(ghive_server *S, uint64_t id, ghive_response *out) {
  if (!S || !out) return GHIVE_ERR_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  if (!S->pending.count(id)) return GHIVE_ERR_INVALID;  // never submitted, or already waited for
  S->cv_done.wait(lk, [&] { return S->done.count(id) != 0 || S->stop; });
  auto it = S->done.find(id);
  if (it == S->done.end()) return GHIVE_ERR_INVALID;  // dropped: the server was destroyed first
  *out = it->second;
  S->done.erase(it);
  S->pending.erase(id);
  return GHIVE_OK;
}

int ghive_server_get_stats(ghive_server *S, ghive_server_stats *out) {
  if (!S || !out) return GHIVE_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  out->completed = S->completed;
  out->rejected = S->rejected;
  out->queued = S->queue.size();
  out->inflight = S->inflight;
  out->inflight_bytes = S->inflight_bytes;
  out->peak_inflight = S->peak_inflight;
  out->peak_inflight_bytes = S->peak_inflight_bytes;
  return GHIVE_OK;
}

ghive_ctx *ghive_server_slot_context(ghive_server *S, uint32_t slot) {
  if (!S || slot >= S->slots.size()) return nullptr;
  return S->slots[slot];
}

void ghive_server_destroy(ghive_server *S) {
  if (!S) return;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    // let admitted work finish; queued work that never started is dropped
    S->cv_done.wait(lk, [&] { return S->inflight == 0; });
    S->stop = true;
  }
  S->cv_queue.notify_all();
  S->cv_slot.notify_all();
  S->cv_done.notify_all();
  if (S->dispatcher.joinable()) S->dispatcher.join();
  for (auto &w : S->workers)
    if (w.joinable()) w.join();
  for (ghive_ctx *c : S->slots) ghive_shutdown(c);
  delete S;
}

}  // extern "C"
