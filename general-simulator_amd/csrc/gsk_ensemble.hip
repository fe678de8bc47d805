// Ensembles of independent instances on one GPU (gsk_ensemble_*).
#include "gsk_internal.h"

// ======================================================================== ensembles
// Many independent instances (any mix of grid sizes and parameters) stepped together.
// Members keep their own buffers; instances with the same grid, variant count and elastic
// data share one B tensor. Three schedules:
//   GSK_ENS_SEQUENTIAL  each member runs its steps alone, one after another;
//   GSK_ENS_STREAMS     members are spread over `branches` streams (greedy balance of
//                       voxels, largest first) and launched eagerly;
//   GSK_ENS_GRAPH       the same branches captured once as a hipGraph holding
//                       `steps_per_graph` steps of every member, replayed as a whole.
// The noise counter of every member lives on the device and is advanced by the last
// kernel of each step, so one graph serves every step.
struct gsk_ensemble {
  std::vector<gsk_sim*> m;
  std::vector<float*> sharedB;
  int dev = 0;
  int mode = GSK_ENS_GRAPH, branches = 0, spg = 10;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipStream_t> bs;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec = nullptr;
  int g_mode = -1, g_branches = -1, g_spg = -1;
  std::vector<unsigned> g_gen;
  // batched schedules: members of one grid size and variant count share every launch
  struct Cls {
    std::vector<int> idx;          // members of the class
    MemberPtrs* dtab = nullptr;    // their device table (buffers and constants)
  };
  std::vector<Cls> cls;
  std::vector<int> loose;          // members that run one by one (other rules, physics or V > 4)
  std::vector<unsigned> cls_gen;   // member generations the tables were built from
};

namespace {

int ens_branches(const gsk_ensemble* e) {
  const int cnt = (int)e->m.size();
  const int b = e->branches > 0 ? e->branches : std::min(cnt, 8);
  return std::max(1, std::min(b, cnt));
}

// members per branch: greedy balance of voxel-fields, largest members first
std::vector<std::vector<int>> ens_assign(const gsk_ensemble* e, int B) {
  std::vector<int> idx(e->m.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  auto work = [&](int i) { return (double)nvox(e->m[i]) * e->m[i]->F; };
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work(a) > work(b); });
  std::vector<double> load(B, 0.0);
  std::vector<std::vector<int>> out(B);
  for (int i : idx) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    out[b].push_back(i);
    load[b] += work(i);
  }
  return out;
}

int ens_streams(gsk_ensemble* e, int B) {
  while ((int)e->bs.size() < B) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    e->bs.push_back(st);
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e->join.push_back(ev);
  }
  return GSK_OK;
}

int sync_all(gsk_ensemble* e) {
  if (e->master) HIPCHK(hipStreamSynchronize(e->master));
  for (hipStream_t st : e->bs) HIPCHK(hipStreamSynchronize(st));
  for (gsk_sim* s : e->m) CHK(sync(s));
  return GSK_OK;
}

int ens_prepare(gsk_ensemble* e) {
  CHK(ensure_device(e->dev));
  for (gsk_sim* s : e->m) {
    if (s->prof) return fail(GSK_ERR_STATE, "an ensemble member is being profiled");
    if (s->cfg.path != GSK_PATH_FUSED32) return fail(GSK_ERR_UNSUPPORTED, "ensembles run the FP32 fused path");
    CHK(prepare(s));
  }
  for (gsk_sim* s : e->m) CHK(sync(s));
  return GSK_OK;
}

// Enqueue `steps` steps of every member on the branch streams, forked from and joined
// back into the master stream (eagerly, or into a graph when the master is capturing).
int ens_launch(gsk_ensemble* e, const std::vector<std::vector<int>>& asg, int steps) {
  const int B = (int)asg.size();
  HIPCHK(hipEventRecord(e->fork, e->master));
  for (int b = 0; b < B; ++b) HIPCHK(hipStreamWaitEvent(e->bs[b], e->fork, 0));
  for (int b = 0; b < B; ++b)
    for (int i : asg[b]) {
      gsk_sim* s = e->m[i];
      hipStream_t own = s->stream;
      s->stream = e->bs[b];
      int rc = GSK_OK;
      for (int k = 0; k < steps && rc == GSK_OK; ++k) rc = dispatch_n<DStep>(s->n, s);
      s->stream = own;
      CHK(rc);
    }
  for (int b = 0; b < B; ++b) {
    HIPCHK(hipEventRecord(e->join[b], e->bs[b]));
    HIPCHK(hipStreamWaitEvent(e->master, e->join[b], 0));
  }
  return GSK_OK;
}

void ens_drop_graph(gsk_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  e->exec = nullptr;
  e->g_mode = e->g_branches = e->g_spg = -1;
  e->g_gen.clear();
}

// Largest grid edge batched into shared launches. Small grids are launch-bound, so one launch
// for a whole size class pays; larger members run better interleaved on the branch streams
// (96 members of 64/128/192: batching 64 only 16.2 ms per ensemble step, also 128 16.7 ms, every
// size 17.3 ms, streams alone 16.6 ms).
#ifndef GSK_ENS_BATCH_MAXN
#define GSK_ENS_BATCH_MAXN 96
#endif

bool batchable(const gsk_sim* s) {
  // (split-y members keep their state and B tensor in another tile order than the batched kernels read)
  return s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_MART && s->cfg.rule == GSK_RULE_SPECTRAL &&
         s->V >= 1 && s->V <= 4 && !s->split;
}

void ens_drop_classes(gsk_ensemble* e) {
  for (auto& c : e->cls)
    if (c.dtab) (void)hipFree(c.dtab);
  e->cls.clear();
  e->loose.clear();
  e->cls_gen.clear();
}

// Size classes of the batched schedules and their device tables, rebuilt when a member's
// constants change (the tables hold each member's buffers and physics constants).
int ens_classes(gsk_ensemble* e) {
  std::vector<unsigned> gen;
  for (gsk_sim* s : e->m) gen.push_back(s->gen);
  if (!e->cls_gen.empty() && gen == e->cls_gen) return GSK_OK;
  ens_drop_graph(e);  // a captured graph refers to the old tables
  CHK(sync_all(e));
  ens_drop_classes(e);
  std::vector<std::pair<int, int>> keys;  // (n, V) of each class
  for (int i = 0; i < (int)e->m.size(); ++i) {
    const gsk_sim* s = e->m[i];
    if (!batchable(s) || s->n > GSK_ENS_BATCH_MAXN) {
      e->loose.push_back(i);
      continue;
    }
    size_t c = 0;
    while (c < keys.size() && keys[c] != std::make_pair(s->n, s->V)) ++c;
    if (c == keys.size()) {
      keys.push_back({s->n, s->V});
      e->cls.emplace_back();
    }
    e->cls[c].idx.push_back(i);
  }
  for (auto& c : e->cls) {
    std::vector<MemberPtrs> h;
    for (int i : c.idx) {
      gsk_sim* s = e->m[i];
      h.push_back(MemberPtrs{s->kbuf, s->state, s->fld, s->Bt, s->dstep, mart_consts(s)});
    }
    HIPCHK(hipMalloc((void**)&c.dtab, h.size() * sizeof(MemberPtrs)));
    HIPCHK(hipMemcpy(c.dtab, h.data(), h.size() * sizeof(MemberPtrs), hipMemcpyHostToDevice));
  }
  e->cls_gen = gen;
  return GSK_OK;
}

// Enqueue `steps` batched steps: one branch stream per size class (every launch covers the
// class's members), the other members spread over further branches as in ens_launch.
int ens_launch_batched(gsk_ensemble* e, int steps) {
  const int C = (int)e->cls.size();
  const int Lb = e->loose.empty() ? 0 : std::min((int)e->loose.size(), ens_branches(e));
  CHK(ens_streams(e, C + Lb));
  HIPCHK(hipEventRecord(e->fork, e->master));
  for (int b = 0; b < C + Lb; ++b) HIPCHK(hipStreamWaitEvent(e->bs[b], e->fork, 0));
  for (int c = 0; c < C; ++c) {
    gsk_sim* s0 = e->m[e->cls[c].idx[0]];
    const int cnt = (int)e->cls[c].idx.size();
    for (int k = 0; k < steps; ++k) CHK(dispatch_n<DBatch>(s0->n, s0, cnt, (const MemberPtrs*)e->cls[c].dtab, e->bs[c]));
  }
  for (int j = 0; j < (int)e->loose.size(); ++j) {
    gsk_sim* s = e->m[e->loose[j]];
    hipStream_t own = s->stream;
    s->stream = e->bs[C + j % Lb];
    int rc = GSK_OK;
    for (int k = 0; k < steps && rc == GSK_OK; ++k) rc = dispatch_n<DStep>(s->n, s);
    s->stream = own;
    CHK(rc);
  }
  for (int b = 0; b < C + Lb; ++b) {
    HIPCHK(hipEventRecord(e->join[b], e->bs[b]));
    HIPCHK(hipStreamWaitEvent(e->master, e->join[b], 0));
  }
  return GSK_OK;
}

bool batched_mode(int mode) { return mode == GSK_ENS_BATCHED || mode == GSK_ENS_BATCHED_GRAPH; }

int ens_graph(gsk_ensemble* e, const std::vector<std::vector<int>>& asg, int spg) {
  std::vector<unsigned> gen;
  for (gsk_sim* s : e->m) gen.push_back(s->gen);
  if (e->exec && e->g_mode == e->mode && e->g_branches == (int)asg.size() && e->g_spg == spg && gen == e->g_gen)
    return GSK_OK;
  ens_drop_graph(e);
  HIPCHK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
  const int rc = batched_mode(e->mode) ? ens_launch_batched(e, spg) : ens_launch(e, asg, spg);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->master, &graph);
  if (rc != GSK_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != GSK_OK ? rc : fail(GSK_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
  }
  const hipError_t ei = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    e->exec = nullptr;
    return fail(GSK_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
  }
  e->g_mode = e->mode;
  e->g_branches = (int)asg.size();
  e->g_spg = spg;
  e->g_gen = gen;
  return GSK_OK;
}

int ens_run(gsk_ensemble* e, int nsteps) {
  CHK(ens_prepare(e));
  if (nsteps == 0) return GSK_OK;
  if (e->mode == GSK_ENS_SEQUENTIAL) {
    for (gsk_sim* s : e->m) {
      for (int k = 0; k < nsteps; ++k) CHK(dispatch_n<DStep>(s->n, s));
      CHK(sync(s));
    }
  } else {
    const bool bat = batched_mode(e->mode);
    const int B = ens_branches(e);
    CHK(ens_streams(e, B));
    if (bat) CHK(ens_classes(e));
    const auto asg = ens_assign(e, B);
    int left = nsteps;
    if ((e->mode == GSK_ENS_GRAPH || e->mode == GSK_ENS_BATCHED_GRAPH) && nsteps >= e->spg) {
      CHK(ens_graph(e, asg, e->spg));
      for (; left >= e->spg; left -= e->spg) HIPCHK(hipGraphLaunch(e->exec, e->master));
    }
    if (left > 0) CHK(bat ? ens_launch_batched(e, left) : ens_launch(e, asg, left));
    HIPCHK(hipStreamSynchronize(e->master));
  }
  for (gsk_sim* s : e->m) s->step += nsteps;
  return GSK_OK;
}

// device copies of the evolving state of every member (real fields, k-space state, step)
struct EnsSnapshot {
  std::vector<float*> fld;
  std::vector<float2*> st;
  std::vector<long long> step;
  ~EnsSnapshot() {
    for (float* p : fld) (void)hipFree(p);
    for (float2* p : st) (void)hipFree(p);
  }
};

int ens_save(gsk_ensemble* e, EnsSnapshot& snap) {
  for (gsk_sim* s : e->m) {
    float* f = nullptr;
    float2* t = nullptr;
    HIPCHK(hipMalloc((void**)&f, (size_t)s->F * nvox(s) * sizeof(float)));
    snap.fld.push_back(f);
    HIPCHK(hipMalloc((void**)&t, state_count(s) * sizeof(float2)));
    snap.st.push_back(t);
    snap.step.push_back(s->step);
    HIPCHK(hipMemcpyAsync(f, s->fld, (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(t, s->state, state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    CHK(sync(s));
  }
  return GSK_OK;
}

int ens_restore(gsk_ensemble* e, const EnsSnapshot& snap) {
  for (size_t i = 0; i < e->m.size(); ++i) {
    gsk_sim* s = e->m[i];
    CHK(sync(s));
    HIPCHK(hipMemcpyAsync(s->fld, snap.fld[i], (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice,
                          s->stream));
    HIPCHK(hipMemcpyAsync(s->state, snap.st[i], state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    const uint32_t v = (uint32_t)snap.step[i];
    HIPCHK(hipMemcpyAsync(s->dstep, &v, sizeof(v), hipMemcpyHostToDevice, s->stream));
    CHK(sync(s));
    s->step = snap.step[i];
    s->sq_ready = false;  // restored fields: the next step transforms eta^2 afresh
  }
  return GSK_OK;
}

}  // namespace

extern "C" {

int gsk_ensemble_create(const gsk_config* cfgs, int count, gsk_ensemble** out) {
  if (!cfgs || !out || count < 1) return fail(GSK_ERR_ARG, "bad argument");
  *out = nullptr;
  for (int i = 1; i < count; ++i)
    if (cfgs[i].device != cfgs[0].device) return fail(GSK_ERR_ARG, "ensemble members must share one device");
  for (int i = 0; i < count; ++i)
    if (box_grid(cfgs[i])) return fail(GSK_ERR_UNSUPPORTED, "ensemble members are cubes");
  CHK(ensure_device(cfgs[0].device));
  gsk_ensemble* e = new gsk_ensemble();
  e->dev = cfgs[0].device;
  auto bail = [&](int rc) { gsk_ensemble_destroy(e); return rc; };
  if (hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
    return bail(fail(GSK_ERR_HIP, "stream/event creation failed"));
  for (int i = 0; i < count; ++i) {
    if (cfgs[i].path != GSK_PATH_FUSED32) return bail(fail(GSK_ERR_UNSUPPORTED, "ensembles run the FP32 fused path"));
    gsk_sim* s = nullptr;
    const int rc = gsk_create(&cfgs[i], &s);
    if (rc != GSK_OK) return bail(rc);
    e->m.push_back(s);
  }
  *out = e;
  return GSK_OK;
}

int gsk_ensemble_destroy(gsk_ensemble* e) {
  if (!e) return GSK_OK;
  (void)hipSetDevice(e->dev);
  if (e->master) (void)hipStreamSynchronize(e->master);
  ens_drop_graph(e);
  ens_drop_classes(e);
  for (gsk_sim* s : e->m) gsk_destroy(s);
  for (float* p : e->sharedB) (void)hipFree(p);
  for (hipStream_t st : e->bs) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->join) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->master) (void)hipStreamDestroy(e->master);
  delete e;
  return GSK_OK;
}

int gsk_ensemble_size(const gsk_ensemble* e) { return e ? (int)e->m.size() : -1; }

int gsk_ensemble_member(gsk_ensemble* e, int i, gsk_sim** out) {
  if (!e || !out || i < 0 || i >= (int)e->m.size()) return fail(GSK_ERR_ARG, "bad member index");
  *out = e->m[i];
  return GSK_OK;
}

int gsk_ensemble_build_btensor(gsk_ensemble* e, const double* C81, const double* eps, int neps) {
  if (!e || !C81 || !eps) return fail(GSK_ERR_ARG, "null argument");
  // one tensor per (grid, variant count, physics, spacing, wave-number convention)
  std::vector<gsk_sim*> owner;
  for (gsk_sim* s : e->m) {
    gsk_sim* o = nullptr;
    for (gsk_sim* c : owner)
      if (c->n == s->n && c->V == s->V && c->cfg.physics == s->cfg.physics && c->cfg.dx == s->cfg.dx &&
          c->cfg.kconv == s->cfg.kconv && c->g.W == s->g.W)
        o = c;
    if (!o) {
      CHK(gsk_build_btensor(s, C81, eps, neps));
      if (!s->sharedB) {  // the ensemble takes ownership of this member's tensor
        e->sharedB.push_back(s->Bt);
        s->sharedB = true;
      }
      owner.push_back(s);
      continue;
    }
    if (s->Bt && !s->sharedB) (void)hipFree(s->Bt);
    s->Bt = o->Bt;
    s->sharedB = true;
    s->cee0 = o->cee0;
    s->bargs = o->bargs;
    s->haveBargs = o->haveBargs;
    s->haveB = true;
    s->dirty = true;
    ++s->gen;
  }
  return GSK_OK;
}

int gsk_ensemble_set_schedule(gsk_ensemble* e, int mode, int branches, int steps_per_graph) {
  if (!e) return fail(GSK_ERR_ARG, "null handle");
  if (mode < GSK_ENS_SEQUENTIAL || mode > GSK_ENS_BATCHED_GRAPH) return fail(GSK_ERR_ARG, "unknown schedule");
  if (branches < 0 || steps_per_graph < 1) return fail(GSK_ERR_ARG, "bad schedule parameters");
  e->mode = mode;
  e->branches = branches;
  e->spg = steps_per_graph;
  return GSK_OK;
}

int gsk_ensemble_get_schedule(const gsk_ensemble* e, int* mode, int* branches, int* steps_per_graph) {
  if (!e || !mode || !branches || !steps_per_graph) return fail(GSK_ERR_ARG, "null argument");
  *mode = e->mode;
  *branches = ens_branches(e);
  *steps_per_graph = e->spg;
  return GSK_OK;
}

int gsk_ensemble_step(gsk_ensemble* e, int nsteps) {
  if (!e) return fail(GSK_ERR_ARG, "null handle");
  if (nsteps < 0) return fail(GSK_ERR_ARG, "negative step count");
  return ens_run(e, nsteps);
}

int gsk_ensemble_autotune(gsk_ensemble* e, int trial_steps, double* table, int max_rows, int* rows) {
  if (!e || trial_steps < 1) return fail(GSK_ERR_ARG, "bad argument");
  CHK(ens_prepare(e));
  EnsSnapshot snap;
  CHK(ens_save(e, snap));
  struct Cand { int mode, branches, spg; double ms; };
  std::vector<Cand> cand;
  const int cnt = (int)e->m.size();
  cand.push_back({GSK_ENS_SEQUENTIAL, 1, 1, 0.0});
  for (int b : {1, 2, 4, 8, 16, 32})
    if (b <= cnt) {
      cand.push_back({GSK_ENS_STREAMS, b, 1, 0.0});
      for (int k : {1, 5, 10})
        if (k <= trial_steps) cand.push_back({GSK_ENS_GRAPH, b, k, 0.0});
    }
  // batched launches per small-grid size class; the branches serve the other members
  for (int b : {8, 16, 32})
    if (b <= cnt || b == 8) {
      const int bb = std::min(b, cnt);
      cand.push_back({GSK_ENS_BATCHED, bb, 1, 0.0});
      for (int k : {1, 5, 10})
        if (k <= trial_steps) cand.push_back({GSK_ENS_BATCHED_GRAPH, bb, k, 0.0});
    }
  const int mode0 = e->mode, b0 = e->branches, k0 = e->spg;
  int rc = GSK_OK;
  for (Cand& c : cand) {
    e->mode = c.mode;
    e->branches = c.branches;
    e->spg = c.spg;
    rc = ens_run(e, c.spg);  // warm-up (and graph build)
    const auto t0 = std::chrono::steady_clock::now();
    if (rc == GSK_OK) rc = ens_run(e, trial_steps);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != GSK_OK) break;
    c.ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial_steps;
  }
  const int rr = ens_restore(e, snap);
  if (rc != GSK_OK || rr != GSK_OK) {
    e->mode = mode0; e->branches = b0; e->spg = k0;
    return rc != GSK_OK ? rc : rr;
  }
  const Cand* best = &cand[0];
  for (const Cand& c : cand)
    if (c.ms < best->ms) best = &c;
  e->mode = best->mode;
  e->branches = best->branches;
  e->spg = best->spg;
  if (rows) *rows = (int)cand.size();
  if (table)
    for (int i = 0; i < (int)cand.size() && i < max_rows; ++i) {
      table[4 * i + 0] = cand[i].mode;
      table[4 * i + 1] = cand[i].branches;
      table[4 * i + 2] = cand[i].spg;
      table[4 * i + 3] = cand[i].ms;
    }
  return GSK_OK;
}

}  // extern "C"
