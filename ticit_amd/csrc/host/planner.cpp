// Learned planner: features, the MLP's training/inference entry points and
// the layout decision made at program creation (planner_mode = 1).
//
// The decision it replaces is ticit's should_select_component_plan
// (src/component_plan.rs:282-298), the dense-versus-components choice that
// the plan blob carries as use_active_components. The learned planner
// predicts, from 12 program features, which layout runs faster on this
// device; when its softmax confidence is within `margin` of 1/2, or no
// model is loaded, the heuristic's choice stands.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "component_plan.hpp"
#include "handle.hpp"

namespace ticit {

struct MlpStepArgs {
  float* params;
  float* m;
  float* v;
  const float* x;
  const int32_t* y;
  const float* t;
  const uint32_t* order;
  uint32_t steps, batch;
  float lr, beta1, beta2, eps, time_weight;
  uint32_t step0;
  int32_t update;
  float* grads;
  float* losses;
};
int mlp_param_count();
int mlp_max_batch();
hipError_t launch_mlp_train(const MlpStepArgs& a, hipStream_t stream);
hipError_t launch_mlp_infer(const float* params, const float* x, uint32_t n, float* out, hipStream_t stream);

// The 12 model features of a program (raw; the model standardizes them).
void planner_features(const Program& p, const ComponentPlan& plan, float out[TICIT_HIP_PLANNER_FEATURES]) {
  int nrot = 0, nprom = 0;
  double mean_dim = 0.0;
  int nq = 0, k = p.initial_k;
  for (const Instr& in : p.instrs) {
    if (in.op == Op::Rotate || in.op == Op::Promote || in.op == Op::Measure) {
      mean_dim += std::ldexp(1.0, k);
      ++nq;
    }
    if (in.op == Op::Rotate) ++nrot;
    if (in.op == Op::Promote) {
      ++nprom;
      ++k;
    }
    if (in.op == Op::Measure) --k;
  }
  mean_dim = nq ? mean_dim / nq : 1.0;
  const double dense_peak = std::ldexp(1.0, p.max_k);
  auto lg = [](double v) { return float(std::log2(1.0 + v)); };
  out[0] = float(p.n);                                  // qubits
  out[1] = float(nrot + nprom);                         // T count (non-Clifford rotations)
  out[2] = float(std::log2(mean_dim));                  // active size: mean log2 dimension
  out[3] = plan.planned ? float(double(plan.peak_live) / dense_peak) : 1.f;  // nonzero fraction
  out[4] = float(p.circuit_layers);                     // Clifford depth
  out[5] = float(p.nrecords);                           // measurement count
  out[6] = lg(plan.dense_work);                         // estimated_dense_work
  out[7] = lg(plan.component_work);                     // estimated_component_work
  out[8] = float(p.max_k);                              // log2 dense_peak_dimension
  out[9] = lg(double(plan.peak_live));                  // component_peak_live_dimension
  out[10] = lg(double(plan.allocated));                 // component_allocated_dimension
  out[11] = float(plan.quantum_ops);                    // quantum-op count
}

namespace {

struct Learned {
  bool set = false;
  std::vector<float> params, mean, scale;
  float margin = 0.f;
  std::map<int, float*> dev_params;  // per device
};
std::mutex g_mu;
Learned g_learned;

}  // namespace

// Decides the layout of a planned program. Returns false when no decision
// was made by the model (no model, unplanned program, low confidence or a
// device error): the caller keeps the heuristic's choice.
bool learned_layout(const Program& p, int device, bool* use_components, float* p_components) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_learned.set) return false;
  ComponentPlan plan = plan_components(p);
  if (!plan.planned) return false;
  float f[TICIT_HIP_PLANNER_FEATURES];
  planner_features(p, plan, f);
  for (int i = 0; i < TICIT_HIP_PLANNER_FEATURES; ++i) f[i] = (f[i] - g_learned.mean[size_t(i)]) / g_learned.scale[size_t(i)];
  float*& dp = g_learned.dev_params[device];
  hipError_t e = hipSuccess;
  if (!dp) {
    e = hipMalloc(reinterpret_cast<void**>(&dp), g_learned.params.size() * sizeof(float));
    if (e == hipSuccess)
      e = hipMemcpy(dp, g_learned.params.data(), g_learned.params.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(dp);
      dp = nullptr;
      return false;
    }
  }
  float* dx = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&dx), 16 * sizeof(float)) != hipSuccess) return false;
  float out[3] = {0, 0, 0};
  e = hipMemcpy(dx, f, sizeof f, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_mlp_infer(dp, dx, 1, dx + TICIT_HIP_PLANNER_FEATURES, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out, dx + TICIT_HIP_PLANNER_FEATURES, sizeof out, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  if (e != hipSuccess) return false;
  const float mx = std::fmax(out[0], out[1]);
  const float e0 = std::exp(out[0] - mx), e1 = std::exp(out[1] - mx);
  const float pc = e1 / (e0 + e1);
  *p_components = pc;
  if (std::fabs(pc - 0.5f) < g_learned.margin) return false;
  *use_components = pc > 0.5f;
  return true;
}

}  // namespace ticit

using namespace ticit;

struct ticit_hip_mlp {
  float* params = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  uint32_t adam_steps = 0;
  ~ticit_hip_mlp() {
    (void)hipFree(params);
    (void)hipFree(m);
    (void)hipFree(v);
  }
};

extern "C" {

ticit_hip_status ticit_hip_planner_features(const uint64_t* blob, size_t nwords, float* out) {
  if (!blob || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  Program p;
  std::string err;
  if (!decode_program(blob, nwords, p, err)) return fail(TICIT_HIP_INVALID_INPUT, err);
  try {
    planner_features(p, plan_components(p), out);
  } catch (const TicitError& e) {
    return fail(ticit_hip_status(e.kind), e.what());
  }
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_mlp_create(const float* params, ticit_hip_mlp** out) {
  if (!params || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  auto* m = new ticit_hip_mlp();
  const size_t bytes = size_t(mlp_param_count()) * sizeof(float);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&m->params), bytes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&m->m), bytes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&m->v), bytes);
  if (e == hipSuccess) e = hipMemcpy(m->params, params, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->m, 0, bytes);
  if (e == hipSuccess) e = hipMemset(m->v, 0, bytes);
  if (e != hipSuccess) {
    delete m;
    return from_hip(e, "allocating the model");
  }
  *out = m;
  return TICIT_HIP_OK;
}

void ticit_hip_mlp_destroy(ticit_hip_mlp* m) { delete m; }

ticit_hip_status ticit_hip_mlp_params_get(const ticit_hip_mlp* m, float* out) {
  if (!m || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  hipError_t e = hipMemcpy(out, m->params, size_t(mlp_param_count()) * sizeof(float), hipMemcpyDeviceToHost);
  return e == hipSuccess ? TICIT_HIP_OK : from_hip(e, "reading the model");
}

namespace {

ticit_hip_status run_train(ticit_hip_mlp* m, const float* x, const int32_t* y, const float* t, size_t n,
                           const uint32_t* order, size_t steps, uint32_t batch, float lr, float time_weight,
                           int update, float* grads, float* losses) {
  if (!m || !x || !y || !t || !order || !losses || (!update && !grads))
    return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  if (batch == 0 || batch > uint32_t(mlp_max_batch()))
    return fail(TICIT_HIP_INVALID_INPUT, "minibatch size must be 1.." + std::to_string(mlp_max_batch()));
  for (size_t i = 0; i < steps * batch; ++i)
    if (order[i] >= n) return fail(TICIT_HIP_INVALID_INPUT, "minibatch index out of range");
  std::vector<void*> bufs;
  auto release = [&]() {
    for (void* p : bufs) (void)hipFree(p);
  };
  auto up = [&](const void* src, size_t bytes, void** dst) {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) return e;
    bufs.push_back(*dst);
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  void *dx, *dy, *dt, *dord, *dloss = nullptr, *dg = nullptr;
  hipError_t e = up(x, n * TICIT_HIP_PLANNER_FEATURES * sizeof(float), &dx);
  if (e == hipSuccess) e = up(y, n * sizeof(int32_t), &dy);
  if (e == hipSuccess) e = up(t, n * sizeof(float), &dt);
  if (e == hipSuccess) e = up(order, steps * batch * sizeof(uint32_t), &dord);
  if (e == hipSuccess && (e = hipMalloc(&dloss, std::max<size_t>(steps, 1) * sizeof(float))) == hipSuccess)
    bufs.push_back(dloss);
  if (e == hipSuccess && !update && (e = hipMalloc(&dg, size_t(mlp_param_count()) * sizeof(float))) == hipSuccess)
    bufs.push_back(dg);
  if (e != hipSuccess) {
    release();
    return from_hip(e, "uploading training data");
  }
  MlpStepArgs a{};
  a.params = m->params;
  a.m = m->m;
  a.v = m->v;
  a.x = static_cast<const float*>(dx);
  a.y = static_cast<const int32_t*>(dy);
  a.t = static_cast<const float*>(dt);
  a.order = static_cast<const uint32_t*>(dord);
  a.steps = uint32_t(update ? steps : 1);
  a.batch = batch;
  a.lr = lr;
  a.beta1 = 0.9f;
  a.beta2 = 0.999f;
  a.eps = 1e-8f;
  a.time_weight = time_weight;
  a.step0 = m->adam_steps;
  a.update = update;
  a.grads = static_cast<float*>(dg);
  a.losses = static_cast<float*>(dloss);
  e = launch_mlp_train(a, nullptr);
  if (e == hipSuccess) e = hipMemcpy(losses, dloss, size_t(a.steps) * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && !update)
    e = hipMemcpy(grads, dg, size_t(mlp_param_count()) * sizeof(float), hipMemcpyDeviceToHost);
  release();
  if (e != hipSuccess) return from_hip(e, "training the model");
  if (update) m->adam_steps += uint32_t(steps);
  return TICIT_HIP_OK;
}

}  // namespace

ticit_hip_status ticit_hip_mlp_train(ticit_hip_mlp* m, const float* x, const int32_t* y, const float* t, size_t n,
                                     const uint32_t* order, size_t steps, uint32_t batch, float lr,
                                     float time_weight, float* losses) {
  return run_train(m, x, y, t, n, order, steps, batch, lr, time_weight, 1, nullptr, losses);
}

ticit_hip_status ticit_hip_mlp_gradients(ticit_hip_mlp* m, const float* x, const int32_t* y, const float* t,
                                         size_t n, float time_weight, float* grads, float* loss) {
  if (n == 0 || n > size_t(mlp_max_batch()))
    return fail(TICIT_HIP_INVALID_INPUT, "gradients take one minibatch of 1.." + std::to_string(mlp_max_batch()));
  std::vector<uint32_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = uint32_t(i);
  return run_train(m, x, y, t, n, order.data(), 1, uint32_t(n), 0.f, time_weight, 0, grads, loss);
}

ticit_hip_status ticit_hip_mlp_predict(const ticit_hip_mlp* m, const float* x, size_t n, float* out) {
  if (!m || (n && (!x || !out))) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  if (n == 0) return TICIT_HIP_OK;
  float *dx = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&dx), n * TICIT_HIP_PLANNER_FEATURES * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dout), n * 3 * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dx, x, n * TICIT_HIP_PLANNER_FEATURES * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_mlp_infer(m->params, dx, uint32_t(n), dout, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out, dout, n * 3 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dout);
  return e == hipSuccess ? TICIT_HIP_OK : from_hip(e, "running the model");
}

ticit_hip_status ticit_hip_planner_set(const float* params, const float* feature_mean, const float* feature_scale,
                                       float margin) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto& kv : g_learned.dev_params) (void)hipFree(kv.second);
  g_learned = Learned{};
  if (!params) return TICIT_HIP_OK;  // unloads the model
  if (!feature_mean || !feature_scale) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  for (int i = 0; i < TICIT_HIP_PLANNER_FEATURES; ++i)
    if (!(feature_scale[i] > 0.f)) return fail(TICIT_HIP_INVALID_INPUT, "feature scales must be positive");
  g_learned.params.assign(params, params + mlp_param_count());
  g_learned.mean.assign(feature_mean, feature_mean + TICIT_HIP_PLANNER_FEATURES);
  g_learned.scale.assign(feature_scale, feature_scale + TICIT_HIP_PLANNER_FEATURES);
  g_learned.margin = margin;
  g_learned.set = true;
  return TICIT_HIP_OK;
}

}  // extern "C"
