// gp_server.cpp — continuous-batching request loop over HBM-resident pools.
//
// Round structure (gp_server_run):
//   1. requests whose arrival time has passed join the ready queue;
//   2. ready requests are admitted into free slots of a pool keyed by
//      (topology signature, dt); a pool is created only while the device
//      memory budget allows it (memory-based admission); all admissions of a
//      round go to the device in one pinned host->device copy and one admit
//      launch per pool;
//   3. one grouped launch advances every busy slot of every pool by up to its
//      pool's steps for the round (a slot runs what is left of its horizon;
//      the kernel counts its remaining steps down, idle slots exit at once).
//      The slowest busy pool runs chunk_steps; a pool whose step costs less
//      runs proportionally more (up to 32x), so every pool's part of the
//      launch takes about as long and a request's latency follows its own
//      mechanism's cost rather than the costliest one in the mix;
//   4. slots whose horizon ended are packed by one gather launch per pool and
//      copied back in one device->host copy; their requests complete with
//      latency = completion - arrival on the server clock.
// A request's arithmetic does not depend on its slot, its pool mates or the
// round it was admitted in: the same kernel body runs it whatever the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gorilla_amd.h"
#include "gp_internal.h"
#include "gp_runtime.h"
#include "gp_serve.h"

using namespace gp;

namespace {

struct Request {
  const gp_model* model = nullptr;
  int model_index = -1;  // the model's entry in the server's device registry
  std::vector<double> q0, v0, tau;
  int64_t horizon = 0;
  double dt = 0, arrival = 0, done_t = 0;
  int pool = -1, slot = -1;
  int64_t remaining = 0;
  std::vector<double> q, v;
  int fault = 0;
  bool finished = false;
};

struct Pool {
  uint64_t sig = 0;
  double dt = 0;
  bool all_rev = false;
  double cost = 1;  // relative time of one step of one wave (bucket_step_cost)
  int chunk = 0;    // steps per round
  Bucket b;
  std::vector<int64_t> slot_req;  // request id per slot, -1 = free
  std::vector<int> free_slots;
  int busy = 0;
  size_t bytes = 0;
};

constexpr int kMaxChunkScale = 32;

template <class T>
size_t round_up(size_t x) {
  return (x + sizeof(T) - 1) / sizeof(T) * sizeof(T);
}

}  // namespace

struct gp_server {
  gp_server_opts o{};
  std::vector<std::unique_ptr<Pool>> pools;
  std::multimap<double, int64_t> future;  // arrival time -> request id
  std::deque<int64_t> ready;
  std::unordered_map<int64_t, Request> req;
  std::vector<int64_t> done;
  int64_t next_id = 1;
  std::chrono::steady_clock::time_point t0;
  hipStream_t stream = nullptr;
  GroupTable* d_group = nullptr;  // n_tables tables of up to kMaxGroup pools each
  int d_group_cap = 0;
  bool group_dirty = true;
  std::vector<int> group_blocks;  // per table
  char* d_stage = nullptr;
  size_t d_stage_cap = 0;
  char* h_stage = nullptr;
  size_t h_stage_cap = 0;
  // device registry of submitted models (their parameters uploaded once):
  // serial -> index into models / d_models; storage bump-allocated in arenas
  std::unordered_map<uint64_t, int> model_index;
  std::vector<ModelDev> models;
  ModelDev* d_models = nullptr;
  size_t d_models_cap = 0;
  bool models_dirty = false;
  std::vector<char*> arenas;
  size_t arena_used = 0, arena_cap = 0;
  double rounds = 0, env_steps = 0, busy_s = 0, admitted = 0, completed = 0, admit_s = 0, retire_s = 0;
  size_t mem_used = 0, budget = 0;
  int64_t running = 0;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int ensure_stage(gp_server* s, size_t bytes) {
  if (bytes > s->d_stage_cap) {
    if (s->d_stage) (void)hipFree(s->d_stage);
    s->d_stage = nullptr;
    size_t cap = std::max(bytes, 2 * s->d_stage_cap);
    hipError_t e = hipMalloc((void**)&s->d_stage, cap);
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(staging): %s", hipGetErrorString(e));
    s->d_stage_cap = cap;
  }
  if (bytes > s->h_stage_cap) {
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    s->h_stage = nullptr;
    size_t cap = std::max(bytes, 2 * s->h_stage_cap);
    hipError_t e = hipHostMalloc((void**)&s->h_stage, cap, hipHostMallocDefault);
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipHostMalloc(staging): %s", hipGetErrorString(e));
    s->h_stage_cap = cap;
  }
  return GP_OK;
}

// Device memory for a registered model's parameters (256-byte aligned),
// from 4 MiB arenas held until the server is destroyed.
int arena_alloc(gp_server* s, size_t bytes, char** out) {
  bytes = (bytes + 255) / 256 * 256;
  if (s->arenas.empty() || s->arena_used + bytes > s->arena_cap) {
    const size_t cap = std::max(bytes, (size_t)4 << 20);
    char* a = nullptr;
    hipError_t e = hipMalloc((void**)&a, cap);
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(model registry): %s", hipGetErrorString(e));
    s->arenas.push_back(a);
    s->arena_used = 0;
    s->arena_cap = cap;
  }
  *out = s->arenas.back() + s->arena_used;
  s->arena_used += bytes;
  return GP_OK;
}

// The model's index in the device registry, uploading its parameters (joint
// kinds, body, contact and half-space values) the first time it is seen.
int register_model(gp_server* s, const gp_model* m, int* out) {
  auto it = s->model_index.find(m->serial);
  if (it != s->model_index.end()) { *out = it->second; return GP_OK; }
  const size_t kb = sizeof(int) * (size_t)m->n, bb = sizeof(double) * m->bp.size(),
               cb = sizeof(double) * m->cp.size(), hb = sizeof(double) * m->hp.size();
  const size_t o_b = (kb + 255) / 256 * 256, o_c = o_b + (bb + 255) / 256 * 256, o_h = o_c + (cb + 255) / 256 * 256;
  char* d = nullptr;
  int st = arena_alloc(s, o_h + hb + 8, &d);
  if (st) return st;
  std::vector<char> h(o_h + hb + 8, 0);
  std::memcpy(h.data(), m->kind.data(), kb);
  if (bb) std::memcpy(h.data() + o_b, m->bp.data(), bb);
  if (cb) std::memcpy(h.data() + o_c, m->cp.data(), cb);
  if (hb) std::memcpy(h.data() + o_h, m->hp.data(), hb);
  hipError_t e = hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "model upload: %s", hipGetErrorString(e));
  s->models.push_back(ModelDev{(const int*)d, (const double*)(d + o_b), (const double*)(d + o_c),
                               (const double*)(d + o_h)});
  const int idx = (int)s->models.size() - 1;
  s->model_index.emplace(m->serial, idx);
  s->models_dirty = true;
  *out = idx;
  return GP_OK;
}

// The registry's table of device pointers, on the device.
int sync_models(gp_server* s) {
  if (!s->models_dirty) return GP_OK;
  if (s->models.size() > s->d_models_cap) {
    const size_t cap = std::max(s->models.size(), 2 * s->d_models_cap + 64);
    if (s->d_models) (void)hipFree(s->d_models);
    s->d_models = nullptr;
    s->d_models_cap = 0;
    hipError_t e = hipMalloc((void**)&s->d_models, cap * sizeof(ModelDev));
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(model table): %s", hipGetErrorString(e));
    s->d_models_cap = cap;
  }
  hipError_t e = hipMemcpy(s->d_models, s->models.data(), s->models.size() * sizeof(ModelDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "model table upload: %s", hipGetErrorString(e));
  s->models_dirty = false;
  return GP_OK;
}

PoolPtrs pool_ptrs(Pool& p) {
  PoolPtrs x{};
  Bucket& b = p.b;
  x.Bp = b.Bp; x.nb = b.nb; x.nq = b.nq; x.nv = b.nv; x.nc = b.nc; x.nh = b.nh;
  x.kind = b.d_kind; x.bp = b.d_bp; x.bpf = b.d_bpf; x.cp = b.d_cp; x.cpf = b.d_cpf; x.hp = b.d_hp; x.hpf = b.d_hpf;
  x.q = b.d_q; x.v = b.d_v; x.tau = b.d_tau; x.vdot = b.d_vdot; x.fault = b.d_fault; x.active = b.d_active;
  return x;
}

size_t pool_bytes(const gp_model* m, int Bp) {
  const size_t per_lane = (size_t)m->n * (NBF * 12 + 4) + (size_t)m->nc * NCF * 12 + (size_t)m->nh * NHF * 12 +
                          (size_t)std::max(1, m->nq) * 8 + 4 * (size_t)std::max(1, m->nv) * 8 + 8;
  return per_lane * Bp;
}

// A pool with a free slot for (sig, dt), creating one if the budget allows.
int find_pool(gp_server* s, const Request& r, int* out) {
  const gp_model* m = r.model;
  for (size_t i = 0; i < s->pools.size(); ++i) {
    Pool& p = *s->pools[i];
    // the all-revolute kernel variant is part of the key: a pool built for
    // revolute-only chains must not take a prismatic joint
    if (p.sig == m->sig && p.dt == r.dt && p.all_rev == m->all_rev && !p.free_slots.empty()) {
      *out = (int)i;
      return GP_OK;
    }
  }
  const int cap = s->o.slots_per_pool;
  if (cap > soa_max_lanes(*m))
    return fail(GP_ERR_INVALID, "slots_per_pool %d above %lld for this model", cap, (long long)soa_max_lanes(*m));
  const int Bp = (cap + 63) / 64 * 64;
  const size_t bytes = pool_bytes(m, Bp);
  if (s->mem_used + bytes > s->budget) {
    // release idle pools (no busy slot) to make room, oldest first
    for (auto& pp : s->pools) {
      if (s->mem_used + bytes <= s->budget) break;
      Pool& p = *pp;
      if (p.busy == 0 && p.bytes > 0) {
        p.b.release();
        s->mem_used -= p.bytes;
        p.bytes = 0;
        p.sig = 0;
        p.free_slots.clear();
        p.slot_req.clear();
        s->group_dirty = true;
      }
    }
    if (s->mem_used + bytes > s->budget) { *out = -1; return GP_OK; }  // wait for slots to free up
  }
  std::unique_ptr<Pool> p(new Pool());
  p->sig = m->sig;
  p->all_rev = m->all_rev;
  p->dt = r.dt;
  std::vector<const gp_model*> lanes(cap, m);
  int st = bucket_build(p->b, lanes, false);
  if (st) return st;
  p->cost = bucket_step_cost(p->b, s->o.precision, s->o.lanes_per_tree);
  hipError_t e = hipMalloc((void**)&p->b.d_active, sizeof(int) * Bp);
  if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(active): %s", hipGetErrorString(e));
  e = hipMemset(p->b.d_active, 0, sizeof(int) * Bp);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "hipMemset(active): %s", hipGetErrorString(e));
  p->slot_req.assign(cap, -1);
  for (int l = cap - 1; l >= 0; --l) p->free_slots.push_back(l);
  p->bytes = bytes;
  s->mem_used += bytes;
  s->pools.push_back(std::move(p));
  s->group_dirty = true;
  *out = (int)s->pools.size() - 1;
  return GP_OK;
}

// The grouped launch tables: every live pool in one table of up to kMaxGroup
// pools, and as many tables (one launch each) as the pools need.
int rebuild_group(gp_server* s) {
  std::vector<GroupTable> tabs;
  std::vector<int> blocks;
  const int block = s->o.block_size;
  for (auto& pp : s->pools) {
    Pool& p = *pp;
    if (p.bytes == 0) continue;  // released
    if (tabs.empty() || tabs.back().n == kMaxGroup) {
      tabs.emplace_back();
      std::memset(&tabs.back(), 0, sizeof(GroupTable));
      blocks.push_back(0);
    }
    GroupTable& t = tabs.back();
    const int lanes = fast_lanes(p.b.fast_topo, s->o.lanes_per_mech, s->o.lanes_per_tree, p.b.nb);
    BucketArgs a = bucket_args(p.b, lanes, nullptr);
    a.dt = p.dt;
    a.n_steps = p.chunk;
    t.args[t.n] = a;
    t.topo[t.n] = p.b.fast_topo;
    t.block_start[t.n] = blocks.back();
    blocks.back() += (int)(((int64_t)p.b.B * lanes + block - 1) / block);
    ++t.n;
    t.block_start[t.n] = blocks.back();
  }
  if ((int)tabs.size() > s->d_group_cap) {
    if (s->d_group) (void)hipFree(s->d_group);
    s->d_group = nullptr;
    s->d_group_cap = 0;
    hipError_t e = hipMalloc((void**)&s->d_group, sizeof(GroupTable) * tabs.size());
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(group): %s", hipGetErrorString(e));
    s->d_group_cap = (int)tabs.size();
  }
  if (!tabs.empty()) {
    hipError_t e = hipMemcpy(s->d_group, tabs.data(), sizeof(GroupTable) * tabs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(GP_ERR_DEVICE, "hipMemcpy(group): %s", hipGetErrorString(e));
  }
  s->group_blocks = blocks;
  s->group_dirty = false;
  return GP_OK;
}

// Pack the admissions of one round (every pool's) behind a table of them,
// copy the whole in one transfer and write them into their slots with one
// launch.
int admit_pools(gp_server* s, std::map<int, std::vector<int64_t>>& adm) {
  int st = sync_models(s);
  if (st) return st;
  // layout: the table, then per pool: lanes, model, steps (ints) then q, v, tau (doubles)
  struct Off { size_t lanes, model, steps, q, v, tau; int K; };
  std::map<int, Off> offs;
  size_t at = round_up<double>(sizeof(AdmitTable));
  for (auto& kv : adm) {
    const Bucket& b = s->pools[kv.first]->b;
    const size_t K = kv.second.size();
    Off o{};
    o.K = (int)K;
    o.lanes = at; at += round_up<double>(K * sizeof(int));
    o.model = at; at += round_up<double>(K * sizeof(int));
    o.steps = at; at += round_up<double>(K * sizeof(int));
    o.q = at; at += K * std::max(1, b.nq) * sizeof(double);
    o.v = at; at += K * std::max(1, b.nv) * sizeof(double);
    o.tau = at; at += K * std::max(1, b.nv) * sizeof(double);
    offs[kv.first] = o;
  }
  st = ensure_stage(s, at);
  if (st) return st;
  char* h = s->h_stage;
  char* d = s->d_stage;
  AdmitTable* tab = (AdmitTable*)h;
  std::memset(tab, 0, sizeof *tab);
  tab->models = s->d_models;
  int total = 0;
  for (auto& kv : adm) {
    const Off& o = offs[kv.first];
    const Bucket& b = s->pools[kv.first]->b;
    for (int k = 0; k < o.K; ++k) {
      const Request& r = s->req[kv.second[k]];
      ((int*)(h + o.lanes))[k] = r.slot;
      ((int*)(h + o.model))[k] = r.model_index;
      ((int*)(h + o.steps))[k] = (int)r.horizon;
      std::memcpy(h + o.q + (size_t)k * b.nq * 8, r.q0.data(), (size_t)b.nq * 8);
      std::memcpy(h + o.v + (size_t)k * b.nv * 8, r.v0.data(), (size_t)b.nv * 8);
      std::memcpy(h + o.tau + (size_t)k * b.nv * 8, r.tau.data(), (size_t)b.nv * 8);
    }
    tab->start[tab->n] = total;
    tab->p[tab->n] = pool_ptrs(*s->pools[kv.first]);
    tab->x[tab->n] = AdmitBatch{o.K, (const int*)(d + o.lanes), (const int*)(d + o.model), (const int*)(d + o.steps),
                                (const double*)(d + o.q), (const double*)(d + o.v), (const double*)(d + o.tau)};
    total += o.K;
    ++tab->n;
  }
  tab->start[tab->n] = total;
  hipError_t e = hipMemcpyAsync(d, h, at, hipMemcpyHostToDevice, s->stream);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "admission copy: %s", hipGetErrorString(e));
  int err = launch_admit((const AdmitTable*)d, total, s->stream);
  if (err) return fail(GP_ERR_DEVICE, "admit launch: %s", hipGetErrorString((hipError_t)err));
  // the pinned staging is reused next round: wait for the copy
  e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "admission sync: %s", hipGetErrorString(e));
  return GP_OK;
}

// Pack every slot that finished this round (one launch behind a table), copy
// the results back in one transfer and complete their requests.
int retire_pools(gp_server* s, std::map<int, std::vector<int64_t>>& fin) {
  struct Off { size_t lanes, q, v, fault; int K; };
  std::map<int, Off> offs;
  const size_t table_bytes = round_up<double>(sizeof(GatherTable));
  size_t at = table_bytes;
  for (auto& kv : fin) {
    const Bucket& b = s->pools[kv.first]->b;
    const size_t K = kv.second.size();
    Off o{};
    o.K = (int)K;
    o.lanes = at; at += round_up<double>(K * sizeof(int));
    o.q = at; at += K * std::max(1, b.nq) * sizeof(double);
    o.v = at; at += K * std::max(1, b.nv) * sizeof(double);
    o.fault = at; at += round_up<double>(K * sizeof(int));
    offs[kv.first] = o;
  }
  int st = ensure_stage(s, at);
  if (st) return st;
  char* h = s->h_stage;
  char* d = s->d_stage;
  GatherTable* tab = (GatherTable*)h;
  std::memset(tab, 0, sizeof *tab);
  int total = 0;
  for (auto& kv : fin) {
    const Off& o = offs[kv.first];
    for (int k = 0; k < o.K; ++k) ((int*)(h + o.lanes))[k] = s->req[kv.second[k]].slot;
    tab->start[tab->n] = total;
    tab->p[tab->n] = pool_ptrs(*s->pools[kv.first]);
    tab->g[tab->n] = GatherBatch{o.K, (const int*)(d + o.lanes), (double*)(d + o.q), (double*)(d + o.v),
                                 (int*)(d + o.fault)};
    total += o.K;
    ++tab->n;
  }
  tab->start[tab->n] = total;
  hipError_t e = hipMemcpyAsync(d, h, at, hipMemcpyHostToDevice, s->stream);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "retire copy: %s", hipGetErrorString(e));
  int err = launch_gather((const GatherTable*)d, total, s->stream);
  if (err) return fail(GP_ERR_DEVICE, "gather launch: %s", hipGetErrorString((hipError_t)err));
  e = hipMemcpyAsync(h + table_bytes, d + table_bytes, at - table_bytes, hipMemcpyDeviceToHost, s->stream);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "result copy: %s", hipGetErrorString(e));
  e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "round sync: %s", hipGetErrorString(e));
  const double t = s->now();
  for (auto& kv : fin) {
    const Off& o = offs[kv.first];
    Pool& p = *s->pools[kv.first];
    for (int k = 0; k < o.K; ++k) {
      Request& r = s->req[kv.second[k]];
      const int nq = p.b.nq, nv = p.b.nv;
      r.q.assign((const double*)(h + o.q) + (size_t)k * nq, (const double*)(h + o.q) + (size_t)(k + 1) * nq);
      r.v.assign((const double*)(h + o.v) + (size_t)k * nv, (const double*)(h + o.v) + (size_t)(k + 1) * nv);
      r.fault = ((const int*)(h + o.fault))[k];
      r.done_t = t;
      r.finished = true;
      p.slot_req[r.slot] = -1;
      p.free_slots.push_back(r.slot);
      --p.busy;
      --s->running;
      s->completed += 1;
      s->done.push_back(kv.second[k]);
    }
  }
  return GP_OK;
}

// A round's admissions or retirements, kMaxGroup pools (one table, one
// launch) at a time.
template <class F>
int by_tables(std::map<int, std::vector<int64_t>>& all, F&& f) {
  std::map<int, std::vector<int64_t>> part;
  for (auto& kv : all) {
    part.emplace(kv.first, std::move(kv.second));
    if (part.size() == (size_t)kMaxGroup) {
      if (int st = f(part)) return st;
      part.clear();
    }
  }
  return part.empty() ? GP_OK : f(part);
}

int admit_round(gp_server* s, std::map<int, std::vector<int64_t>>& adm) {
  return by_tables(adm, [s](std::map<int, std::vector<int64_t>>& m) { return admit_pools(s, m); });
}

int retire_round(gp_server* s, std::map<int, std::vector<int64_t>>& fin) {
  return by_tables(fin, [s](std::map<int, std::vector<int64_t>>& m) { return retire_pools(s, m); });
}

int run_rounds(gp_server* s, int max_rounds, int* rounds_done) {
  int rounds = 0;
  const int C = s->o.chunk_steps;
  for (;;) {
    if (max_rounds > 0 && rounds >= max_rounds) break;
    const double now = s->now();
    while (!s->future.empty() && s->future.begin()->first <= now) {
      s->ready.push_back(s->future.begin()->second);
      s->future.erase(s->future.begin());
    }
    // admissions: every ready request that finds a slot, in arrival order
    std::map<int, std::vector<int64_t>> adm;
    std::deque<int64_t> still;
    for (int64_t id : s->ready) {
      Request& r = s->req[id];
      int pi = -1;
      int st = find_pool(s, r, &pi);
      if (st) return st;
      if (pi < 0) { still.push_back(id); continue; }
      Pool& p = *s->pools[pi];
      r.pool = pi;
      r.slot = p.free_slots.back();
      p.free_slots.pop_back();
      p.slot_req[r.slot] = id;
      ++p.busy;
      r.remaining = r.horizon;
      adm[pi].push_back(id);
      ++s->running;
      s->admitted += 1;
    }
    s->ready.swap(still);
    if (!adm.empty()) {
      const double t_adm = s->now();
      int st = admit_round(s, adm);
      if (st) return st;
      s->admit_s += s->now() - t_adm;
    }
    if (s->running == 0) {
      if (s->ready.empty() && s->future.empty()) break;
      if (!s->ready.empty()) return fail(GP_ERR_BUSY, "queued requests do not fit the memory budget");
      const double wait = s->future.begin()->first - s->now();
      if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
      continue;
    }
    // steps per pool this round: the costliest busy pool runs C
    double cmax = 0;
    for (auto& pp : s->pools)
      if (pp->busy) cmax = std::max(cmax, pp->cost);
    for (auto& pp : s->pools) {
      const int ch = pp->busy ? std::min(C * kMaxChunkScale, std::max(C, (int)std::lround(C * cmax / pp->cost))) : C;
      if (ch != pp->chunk) { pp->chunk = ch; s->group_dirty = true; }
    }
    if (s->group_dirty) {
      int st = rebuild_group(s);
      if (st) return st;
    }
    const double ta = s->now();
    for (size_t t = 0; t < s->group_blocks.size(); ++t) {
      int e = launch_grouped_step(s->d_group + t, s->group_blocks[t], s->o.precision, 0.0, C, s->o.block_size,
                                  s->stream);
      if (e) return fail(GP_ERR_DEVICE, "grouped launch: %s", hipGetErrorString((hipError_t)e));
    }
    // the round's steps end before its retirements are packed (the stream
    // orders them anyway; the wait keeps the step out of retire_s)
    hipError_t he = hipStreamSynchronize(s->stream);
    if (he != hipSuccess) return fail(GP_ERR_DEVICE, "round sync: %s", hipGetErrorString(he));
    // host mirror of the kernel's count-down
    std::map<int, std::vector<int64_t>> fin;
    for (size_t pi = 0; pi < s->pools.size(); ++pi) {
      Pool& p = *s->pools[pi];
      if (!p.busy) continue;
      for (size_t l = 0; l < p.slot_req.size(); ++l) {
        const int64_t id = p.slot_req[l];
        if (id < 0) continue;
        Request& r = s->req[id];
        const int64_t k = std::min<int64_t>(p.chunk, r.remaining);
        r.remaining -= k;
        s->env_steps += (double)k;
        if (r.remaining == 0) fin[(int)pi].push_back(id);
      }
    }
    if (!fin.empty()) {
      const double t_ret = s->now();
      int st = retire_round(s, fin);
      if (st) return st;
      s->retire_s += s->now() - t_ret;
    }
    s->busy_s += s->now() - ta;
    s->rounds += 1;
    ++rounds;
  }
  if (rounds_done) *rounds_done = rounds;
  return GP_OK;
}

}  // namespace

extern "C" {

int gp_server_create(const gp_server_opts* opts, gp_server** out) {
  if (!out) return fail(GP_ERR_INVALID, "null argument");
  std::unique_ptr<gp_server> s(new (std::nothrow) gp_server());
  if (!s) return fail(GP_ERR_NOMEM, "out of host memory");
  if (opts) s->o = *opts;
  if (s->o.precision != GP_PREC_F64 && s->o.precision != GP_PREC_MIXED)
    return fail(GP_ERR_INVALID, "unknown precision %d", s->o.precision);
  if (s->o.chunk_steps <= 0) s->o.chunk_steps = 100;
  if (s->o.slots_per_pool <= 0) s->o.slots_per_pool = 4096;
  if (s->o.block_size <= 0) s->o.block_size = 128;
  if (s->o.block_size % 64 || s->o.block_size > 256) return fail(GP_ERR_INVALID, "block_size must be 64, 128, 192 or 256");
  if (s->o.mem_budget_bytes < 0) return fail(GP_ERR_INVALID, "negative memory budget");
  if (s->o.lanes_per_mech != 0 && s->o.lanes_per_mech != 1 && s->o.lanes_per_mech != 2 && s->o.lanes_per_mech != 4 &&
      s->o.lanes_per_mech != 8)
    return fail(GP_ERR_INVALID, "lanes_per_mech must be 1, 2, 4 or 8");
  if (s->o.lanes_per_tree != 0 && s->o.lanes_per_tree != 1 && s->o.lanes_per_tree != 4)
    return fail(GP_ERR_INVALID, "lanes_per_tree must be 1 or 4");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(GP_ERR_DEVICE, "no HIP device");
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(GP_ERR_DEVICE, "hipMemGetInfo failed");
  s->budget = s->o.mem_budget_bytes > 0 ? std::min<size_t>((size_t)s->o.mem_budget_bytes, free_b) : free_b / 2;
  if (hipStreamCreate(&s->stream) != hipSuccess) return fail(GP_ERR_DEVICE, "hipStreamCreate failed");
  s->t0 = std::chrono::steady_clock::now();
  *out = s.release();
  return GP_OK;
}

void gp_server_destroy(gp_server* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  s->pools.clear();
  if (s->d_group) (void)hipFree(s->d_group);
  if (s->d_models) (void)hipFree(s->d_models);
  for (char* a : s->arenas) (void)hipFree(a);
  if (s->d_stage) (void)hipFree(s->d_stage);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int gp_server_clock(const gp_server* s, double* now_s) {
  if (!s || !now_s) return fail(GP_ERR_INVALID, "null argument");
  *now_s = s->now();
  return GP_OK;
}

int gp_server_submit(gp_server* s, const gp_model* m, const double* q0, const double* v0, const double* tau,
                     int64_t horizon, double dt, double arrival_s, int64_t* id) {
  if (!s || !m || !q0 || !v0 || !id) return fail(GP_ERR_INVALID, "null argument");
  if (m->fast_topo < 0) return fail(GP_ERR_UNSUPPORTED, "serving needs a topology with a FAST kernel");
  if (horizon < 0 || horizon > (int64_t)1 << 30) return fail(GP_ERR_INVALID, "horizon out of range");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(GP_ERR_INVALID, "dt must be positive");
  try {
    Request r;
    r.model = m;
    int st = register_model(s, m, &r.model_index);
    if (st) return st;
    r.q0.assign(q0, q0 + m->nq);
    r.v0.assign(v0, v0 + m->nv);
    r.tau.assign(m->nv, 0.0);
    if (tau) r.tau.assign(tau, tau + m->nv);
    r.horizon = horizon;
    r.dt = dt;
    r.arrival = arrival_s < 0 ? s->now() : arrival_s;
    const int64_t rid = s->next_id++;
    if (horizon == 0) {  // nothing to run: complete at once
      r.q = r.q0;
      r.v = r.v0;
      r.finished = true;
      r.done_t = std::max(r.arrival, s->now());
      s->req.emplace(rid, std::move(r));
      s->done.push_back(rid);
      s->completed += 1;
    } else {
      s->future.emplace(r.arrival, rid);
      s->req.emplace(rid, std::move(r));
    }
    *id = rid;
    return GP_OK;
  } catch (const std::bad_alloc&) {
    return fail(GP_ERR_NOMEM, "out of host memory");
  }
}

int gp_server_run(gp_server* s, int32_t max_rounds, int32_t* rounds_done) {
  if (!s) return fail(GP_ERR_INVALID, "null server");
  try {
    return run_rounds(s, max_rounds, rounds_done);
  } catch (const std::bad_alloc&) {
    return fail(GP_ERR_NOMEM, "out of host memory");
  }
}

int gp_server_pending(const gp_server* s, int64_t* queued, int64_t* running) {
  if (!s) return fail(GP_ERR_INVALID, "null server");
  if (queued) *queued = (int64_t)(s->future.size() + s->ready.size());
  if (running) *running = s->running;
  return GP_OK;
}

int gp_server_completed(gp_server* s, int64_t* ids, int64_t max_ids, int64_t* n) {
  if (!s || !n) return fail(GP_ERR_INVALID, "null argument");
  const int64_t k = std::min<int64_t>(max_ids, (int64_t)s->done.size());
  for (int64_t i = 0; i < k; ++i) ids[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return GP_OK;
}

int gp_server_result(gp_server* s, int64_t id, double* q, double* v, double* latency_s, int32_t* fault) {
  if (!s) return fail(GP_ERR_INVALID, "null server");
  auto it = s->req.find(id);
  if (it == s->req.end()) return fail(GP_ERR_INVALID, "unknown request %lld", (long long)id);
  Request& r = it->second;
  if (!r.finished) return fail(GP_ERR_BUSY, "request %lld has not completed", (long long)id);
  if (q) std::copy(r.q.begin(), r.q.end(), q);
  if (v) std::copy(r.v.begin(), r.v.end(), v);
  if (latency_s) *latency_s = r.done_t - r.arrival;
  if (fault) *fault = r.fault;
  s->req.erase(it);
  return GP_OK;
}

int gp_server_stats(const gp_server* s, double* st) {
  if (!s || !st) return fail(GP_ERR_INVALID, "null argument");
  st[0] = s->rounds;
  st[1] = s->env_steps;
  st[2] = s->busy_s;
  st[3] = s->admitted;
  st[4] = s->completed;
  st[5] = (double)std::count_if(s->pools.begin(), s->pools.end(), [](const std::unique_ptr<Pool>& p) { return p->bytes > 0; });
  st[6] = (double)s->mem_used;
  st[7] = (double)s->budget;
  st[8] = s->admit_s;
  st[9] = s->retire_s;
  return GP_OK;
}

}  // extern "C"
