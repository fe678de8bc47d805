// Checkpointed reverse mode of the FP32 martensite step (gsk_adjoint_*).
#include "gsk_internal.h"

// ======================================================================== adjoint
// Reverse mode of the FP32 martensite step with checkpointing. gsk_adjoint_forward steps
// the handle from its current state and keeps (real field, k-space state, step) every K
// steps plus the final state; gsk_adjoint_backward walks the segments backwards: it
// restores a checkpoint, recomputes the segment's K states into a buffer (the restored
// state reproduces the original trajectory bit for bit), and runs the reverse steps over
// it. K = 1 keeps every state and recomputes nothing (store-all).
struct AdjWS {
  int T = 0, K = 0;
  long long step_end = 0;
  std::vector<float*> ck_fld;
  std::vector<float2*> ck_state;
  std::vector<long long> ck_step;
  float* fin_fld = nullptr;
  float2* fin_state = nullptr;
  float *seg = nullptr, *lam = nullptr, *mu = nullptr, *rq = nullptr, *bh = nullptr;
  float2 *kb = nullptr, *k2 = nullptr;
  double* part = nullptr;
  std::vector<double> hpart;
  int nb = 512;
  double ms_fwd = 0.0, ms_bwd = 0.0;
  double bytes_ck = 0.0, bytes_seg = 0.0, bytes_work = 0.0;
};

namespace {


// copy the evolving state (real fields, k-space state) of the handle into / out of a slot
int adj_put(gsk_sim* s, float* fld, float2* st) {
  HIPCHK(hipMemcpyAsync(fld, s->fld, (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(st, s->state, state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  return GSK_OK;
}
int adj_get(gsk_sim* s, const float* fld, const float2* st, long long step) {
  HIPCHK(hipMemcpyAsync(s->fld, fld, (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->state, st, state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  const uint32_t v = (uint32_t)step;
  HIPCHK(hipMemcpyAsync(s->dstep, &v, sizeof(v), hipMemcpyHostToDevice, s->stream));
  CHK(sync(s));
  s->step = step;
  s->sq_ready = false;  // new fields: the next step transforms eta^2 afresh
  return GSK_OK;
}

template <typename T>
int dev_alloc(T*& p, size_t count, double& tally) {
  if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess)
    return fail(GSK_ERR_HIP, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  tally += (double)count * sizeof(T);
  return GSK_OK;
}

AdjK adj_consts(const gsk_sim* s, uint32_t step) {
  const MartK m = mart_consts(s);
  AdjK k{};
  k.A = m.A; k.B = m.B; k.C = m.C; k.two_wel = m.two_wel; k.noise = m.noise;
  k.dtL = m.dtL; k.dtLk = m.dtLk; k.kfac = m.kfac; k.inv_n3 = m.inv_n3;
  k.V = s->V;
  k.step = step;
  k.seed = m.seed;
  return k;
}

}  // namespace

void adj_free(AdjWS* w) {
  if (!w) return;
  for (float* p : w->ck_fld) (void)hipFree(p);
  for (float2* p : w->ck_state) (void)hipFree(p);
  free_dev(w->fin_fld); free_dev(w->fin_state);
  free_dev(w->seg); free_dev(w->lam); free_dev(w->mu); free_dev(w->rq); free_dev(w->bh);
  free_dev(w->kb); free_dev(w->k2); free_dev(w->part);
  delete w;
}

extern "C" {

int gsk_adjoint_forward(gsk_sim* s, int nsteps, int K) {
  if (!s || nsteps < 1 || K < 1) return fail(GSK_ERR_ARG, "bad argument");
  if (s->cfg.path != GSK_PATH_FUSED32 || s->cfg.physics != GSK_PHYS_MART || s->cfg.rule != GSK_RULE_SPECTRAL || s->bx)
    return fail(GSK_ERR_UNSUPPORTED, "the adjoint covers the FP32 fused spectral martensite step on cubes");
  if (s->V > 4) return fail(GSK_ERR_UNSUPPORTED, "the adjoint supports up to 4 variants");
  CHK(ensure_device(s->dev));
  CHK(prepare(s));
  adj_free(s->adj);
  s->adj = new AdjWS();
  AdjWS& w = *s->adj;
  w.T = nsteps;
  w.K = K;
  const int nck = (nsteps + K - 1) / K;
  const size_t fcount = (size_t)s->F * nvox(s), scount = state_count(s);
  const auto t0 = std::chrono::steady_clock::now();
  for (int j = 0; j < nck; ++j) {
    float* f = nullptr;
    float2* t = nullptr;
    CHK(dev_alloc(f, fcount, w.bytes_ck));
    w.ck_fld.push_back(f);
    CHK(dev_alloc(t, scount, w.bytes_ck));
    w.ck_state.push_back(t);
    w.ck_step.push_back(s->step);
    CHK(adj_put(s, f, t));
    const int len = std::min(K, nsteps - j * K);
    for (int i = 0; i < len; ++i) {
      CHK(dispatch_n<DStep>(s->n, s));
      ++s->step;
    }
  }
  CHK(dev_alloc(w.fin_fld, fcount, w.bytes_ck));
  CHK(dev_alloc(w.fin_state, scount, w.bytes_ck));
  CHK(adj_put(s, w.fin_fld, w.fin_state));
  CHK(sync(s));
  w.step_end = s->step;
  w.ms_fwd = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GSK_OK;
}

int gsk_adjoint_backward(gsk_sim* s, const float* dL_detaT, float* grad_eta0, double* grad_params) {
  if (!s || !dL_detaT) return fail(GSK_ERR_ARG, "null argument");
  if (!s->adj || s->adj->T == 0) return fail(GSK_ERR_STATE, "no recorded forward run (call gsk_adjoint_forward)");
  if (!s->haveBargs) return fail(GSK_ERR_STATE, "B tensor not built");
  CHK(ensure_device(s->dev));
  AdjWS& w = *s->adj;
  const int V = s->V, K = w.K, T = w.T, nck = (T + K - 1) / K, nb = w.nb;
  const size_t N3 = nvox(s), VN = (size_t)V * N3;
  if (!w.lam) {
    if (K > 1) CHK(dev_alloc(w.seg, (size_t)(K + 1) * VN, w.bytes_seg));
    CHK(dev_alloc(w.lam, VN, w.bytes_work));
    CHK(dev_alloc(w.mu, VN, w.bytes_work));
    CHK(dev_alloc(w.rq, 2 * VN, w.bytes_work));
    CHK(dev_alloc(w.kb, (size_t)(2 * V) * s->g.kfs, w.bytes_work));
    CHK(dev_alloc(w.k2, (size_t)V * s->g.kfs, w.bytes_work));
    CHK(dev_alloc(w.bh, (size_t)(V * (V + 1) / 2) * s->g.kfs, w.bytes_work));
    // FP64 partial sums of every reverse step of a segment, copied to the host once per segment
    CHK(dev_alloc(w.part, (size_t)K * nb * (2 + V * V), w.bytes_work));
    w.hpart.resize((size_t)K * nb * (2 + V * V));
    hipLaunchKernelGGL(k_btensor_half<>, dim3(grid_1d((long)s->n * s->n * (s->n / 2 + 1))), dim3(256), 0, s->stream,
                       w.bh, s->g, s->bargs);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(w.lam, dL_detaT, VN * sizeof(float), hipMemcpyHostToDevice, s->stream));
  const double dt = s->cfg.dt, dtL = s->cfg.dt * s->cfg.L, kap = s->cfg.kappa, wel = s->cfg.w_el;
  const double inv_n3 = 1.0 / (double)N3;
  double gL = 0.0, gk = 0.0;
  std::vector<double> gs(V, 0.0);
  const auto t0 = std::chrono::steady_clock::now();
  for (int j = nck - 1; j >= 0; --j) {
    const int len = std::min(K, T - j * K);
    CHK(adj_get(s, w.ck_fld[j], w.ck_state[j], w.ck_step[j]));
    if (K > 1) {  // recompute the segment's states
      HIPCHK(hipMemcpyAsync(w.seg, s->fld, VN * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
      for (int i = 1; i <= len; ++i) {
        CHK(dispatch_n<DStep>(s->n, s));
        ++s->step;
        HIPCHK(hipMemcpyAsync(w.seg + (size_t)i * VN, s->fld, VN * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
      }
    }
    const size_t stride = (size_t)nb * (2 + V * V);
    for (int i = len - 1; i >= 0; --i) {
      const float* eta = K > 1 ? w.seg + (size_t)i * VN : w.ck_fld[j];
      const float* eta1 = K > 1 ? w.seg + (size_t)(i + 1) * VN : (j + 1 < nck ? w.ck_fld[j + 1] : w.fin_fld);
      const AdjK k = adj_consts(s, (uint32_t)(w.ck_step[j] + i));
      CHK(dispatch_n<DAdj>(s->n, s, eta, eta1, w.lam, w.mu, w.rq, w.kb, w.k2, (const float*)w.bh,
                           w.part + (size_t)i * stride, nb, k));
    }
    HIPCHK(hipMemcpyAsync(w.hpart.data(), w.part, (size_t)len * stride * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
    CHK(sync(s));
    // the step's sums in the same order as the reverse sweep: gradients do not depend on K
    for (int i = len - 1; i >= 0; --i) {
      const double* hp = w.hpart.data() + (size_t)i * stride;
      double X = 0.0, muF = 0.0;
      std::vector<double> G((size_t)V * V, 0.0);
      for (int b = 0; b < nb; ++b) {
        X += hp[b];
        for (int r = 0; r < V * V; ++r) G[r] += hp[nb + (size_t)b * V * V + r];
        muF += hp[(size_t)nb + (size_t)nb * V * V + b];
      }
      X *= inv_n3;
      gL += -dt * kap * X - dt * muF;
      gk += -dtL * X;
      for (int r = 0; r < V; ++r)
        for (int q = 0; q < V; ++q) {
          const double t = -2.0 * dtL * wel * G[r * V + q] * inv_n3;
          gs[r] += t;
          gs[q] += t;
        }
    }
  }
  if (grad_eta0) HIPCHK(hipMemcpyAsync(grad_eta0, w.lam, VN * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  CHK(adj_get(s, w.fin_fld, w.fin_state, w.step_end));
  w.ms_bwd = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (grad_params) {
    for (int v = 0; v < V; ++v) grad_params[v] = gs[v];
    grad_params[V] = gL;
    grad_params[V + 1] = gk;
  }
  return GSK_OK;
}

int gsk_adjoint_info(const gsk_sim* s, double* out) {
  if (!s || !out) return fail(GSK_ERR_ARG, "null argument");
  if (!s->adj) return fail(GSK_ERR_STATE, "no recorded forward run");
  const AdjWS& w = *s->adj;
  out[0] = w.bytes_ck;
  out[1] = w.bytes_seg;
  out[2] = w.bytes_work;
  out[3] = w.ms_fwd;
  out[4] = w.ms_bwd;
  out[5] = (double)w.ck_fld.size() + 1.0;  // checkpoints incl. the final state
  return GSK_OK;
}

int gsk_adjoint_release(gsk_sim* s) {
  if (!s) return fail(GSK_ERR_ARG, "null handle");
  CHK(ensure_device(s->dev));
  adj_free(s->adj);
  s->adj = nullptr;
  return GSK_OK;
}

}  // extern "C"
