// neural.hip -- activation procedures and the FP64 feed-forward network of
// oap's neural package, on top of the procedures API.
//
// Activations (CuProceduresApi.h:239-290; kernels CuSigmoidProcedures.h,
// CuTanhProcedures.h, CuSinProcedures.h, CuFuncProcedures.h): applied to
// every plane that the output has, each plane on its own (not a complex
// function); the dims variant limits the region to x < dims[0], y < dims[1];
// the periodic variant to x < d[0][0], y < rows and (y mod d[1][1]) <
// d[0][1] (cuda_inRangePD).
//
// Network (oapNeuralDevice/oapNetwork.cpp, oapNeuralHost/oapGenericNeuralApi.h
// :195-364): per-sample forward propagation (dotProduct with dims, then the
// layer activation), output error e = y - t (substract), back propagation
// (derivative of the activation in place on the sums, hadamardProductVec,
// transposed weights, dotProduct with dims), nabla W = tensorProduct of the
// transposed inputs and the next layer's errors accumulated with add, and the
// update W <- W - (lr / N) sum nabla W with N the number of accumulated
// errors.  Every step is a call of this library's procedures on the
// network's stream.
#include <cmath>
#include <vector>

#include "internal.h"

using namespace oaph;

namespace {

constexpr int kThreads = 256;

// region selector: 0 all, 1 dims {w, h}, 2 periodic {w, h_in_period, period}
struct Region {
  int mode;
  uint32_t w, h, period;
};

__device__ __forceinline__ double act(int fn, double out_old, double v) {
  switch (fn) {
    case OAPH_FN_SIGMOID: return 1.0 / (1.0 + exp(-v));
    case OAPH_FN_DSIGMOID: {
      const double s = 1.0 / (1.0 + exp(-v));
      return s * (1.0 - s);
    }
    case OAPH_FN_MULTIPLY_DSIGMOID: {
      const double s = 1.0 / (1.0 + exp(-v));
      return out_old * s * (1.0 - s);
    }
    case OAPH_FN_LINEAR: return v;
    case OAPH_FN_DLINEAR: return 1.0;
    case OAPH_FN_TANH: return tanh(v);
    case OAPH_FN_DTANH: {
      const double t = tanh(v);
      return 1.0 - t * t;
    }
    case OAPH_FN_MULTIPLY_DTANH: {
      const double t = tanh(v);
      return out_old * (1.0 - t * t);
    }
    case OAPH_FN_SIN: return sin(v);
    case OAPH_FN_DSIN: return cos(v);
    default: return out_old * cos(v);  // OAPH_FN_MULTIPLY_DSIN
  }
}

__global__ __launch_bounds__(kThreads) void k_activation(int fn, Region rg, uint32_t cols,
                                                         uint32_t rows, double* __restrict__ ore,
                                                         double* __restrict__ oim,
                                                         const double* __restrict__ ire,
                                                         const double* __restrict__ iim) {
  const size_t n = (size_t)cols * rows;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t y = (uint32_t)(i / cols), x = (uint32_t)(i - (size_t)y * cols);
    bool in = true;
    if (rg.mode == 1) in = x < rg.w && y < rg.h;
    else if (rg.mode == 2) in = x < rg.w && (y % rg.period) < rg.h;
    if (!in) continue;
    if (ore) ore[i] = act(fn, ore[i], ire[i]);
    if (oim) oim[i] = act(fn, oim[i], iim[i]);
  }
}

oaph_status launch_activation(oaph_api* api, int fn, oaph_matrix* out, const oaph_matrix* in,
                              Region rg) {
  OAPH_CHECK(api, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  OAPH_CHECK(out && in, OAPH_ERR_ARG);
  OAPH_CHECK(fn >= OAPH_FN_SIGMOID && fn <= OAPH_FN_MULTIPLY_DSIN, OAPH_ERR_ARG);
  // check_isEqualDim (GenericProceduresApi.h:552)
  OAPH_CHECK(out->columns == in->columns && out->rows == in->rows, OAPH_ERR_SHAPE);
  OAPH_CHECK((!out->reValues || in->reValues) && (!out->imValues || in->imValues),
             OAPH_ERR_PLANE);
  OAPH_CHECK(rg.mode != 2 || rg.period > 0, OAPH_ERR_ARG);
  const size_t n = numel(out);
  if (n == 0) return OAPH_OK;
  const unsigned grid = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_activation, dim3(grid), dim3(kThreads), 0, api->stream, fn, rg,
                     out->columns, out->rows, out->reValues, out->imValues, in->reValues,
                     in->imValues);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

// A device column vector / matrix owned by the network (re plane only, as the
// reference's newDeviceReMatrix).
struct Mat {
  oaph_matrix* m = nullptr;
  oaph_status alloc(uint32_t cols, uint32_t rows) {
    return oaph_matrix_new(&m, 1, 0, cols, rows);
  }
  void release() {
    if (m) oaph_matrix_free(m);
    m = nullptr;
  }
};

struct Layer {
  uint32_t neurons = 0, bias = 0;
  int activation = OAPH_FN_TANH;
  Mat inputs, sums, errors, errors_aux, tinputs;
  Mat weights, tweights, weights1, weights2;  // to the next layer: rows = next neurons
  uint32_t total() const { return neurons + bias; }
};

int derivative_of(int fn) {
  switch (fn) {
    case OAPH_FN_SIGMOID: return OAPH_FN_DSIGMOID;
    case OAPH_FN_LINEAR: return OAPH_FN_DLINEAR;
    case OAPH_FN_SIN: return OAPH_FN_DSIN;
    default: return OAPH_FN_DTANH;
  }
}

}  // namespace

struct oaph_network {
  oaph_api* api = nullptr;
  bool own_api = false;
  std::vector<Layer> layers;
  double learning_rate = 0.1;
  Mat expected;
  std::vector<double> errors;  // 0.5 e^2 per accumulated output (MSE bookkeeping)
  std::vector<double> host;    // staging
};

namespace {

oaph_status set_bias(Layer& L) {
  if (!L.bias) return OAPH_OK;
  return oaph_matrix_set_value(L.inputs.m, L.total() - 1, 1.0, 0.0);
}

oaph_status connect(Layer& cur, const Layer& next) {
  const uint32_t c = cur.total(), r = next.neurons;
  OAPH_TRY(cur.weights.alloc(c, r));
  OAPH_TRY(cur.tweights.alloc(r, c));
  OAPH_TRY(cur.weights1.alloc(c, r));
  OAPH_TRY(cur.weights2.alloc(c, r));
  return OAPH_OK;
}

oaph_status activation_dims(oaph_api* api, int fn, oaph_matrix* out, const oaph_matrix* in,
                            uint32_t w, uint32_t h) {
  Region rg{1, w, h, 0};
  return launch_activation(api, fn, out, in, rg);
}

}  // namespace

extern "C" {

oaph_status oaph_activation(oaph_api* api, int fn, oaph_matrix* out, const oaph_matrix* in) {
  return launch_activation(api, fn, out, in, Region{0, 0, 0, 0});
}

oaph_status oaph_activation_dims(oaph_api* api, int fn, oaph_matrix* out, const oaph_matrix* in,
                                 const uint32_t dims[2]) {
  OAPH_CHECK(dims, OAPH_ERR_ARG);
  return launch_activation(api, fn, out, in, Region{1, dims[0], dims[1], 0});
}

oaph_status oaph_activation_dims_periodic(oaph_api* api, int fn, oaph_matrix* out,
                                          const oaph_matrix* in, const uint32_t dims[2][2]) {
  OAPH_CHECK(dims, OAPH_ERR_ARG);
  return launch_activation(api, fn, out, in, Region{2, dims[0][0], dims[0][1], dims[1][1]});
}

oaph_status oaph_network_create(oaph_network** out, oaph_api* api) {
  OAPH_CHECK(out, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  oaph_network* n = new oaph_network();
  if (api) {
    n->api = api;
  } else {
    oaph_status st = oaph_api_create(&n->api);
    if (st != OAPH_OK) {
      delete n;
      return st;
    }
    n->own_api = true;
  }
  *out = n;
  return OAPH_OK;
}

oaph_status oaph_network_destroy(oaph_network* n) {
  if (!n) return OAPH_OK;
  (void)oaph_api_synchronize(n->api);
  for (Layer& L : n->layers) {
    for (Mat* m : {&L.inputs, &L.sums, &L.errors, &L.errors_aux, &L.tinputs, &L.weights,
                   &L.tweights, &L.weights1, &L.weights2})
      m->release();
  }
  n->expected.release();
  if (n->own_api) oaph_api_destroy(n->api);
  delete n;
  return OAPH_OK;
}

// Network::createLayer(neurons, addBias, activation) (oapNetwork.h:53-54);
// the previous layer is connected to the new one.
oaph_status oaph_network_add_layer(oaph_network* n, uint32_t neurons, int has_bias,
                                   int activation) {
  OAPH_CHECK(n && neurons > 0, OAPH_ERR_ARG);
  OAPH_CHECK(activation == OAPH_FN_SIGMOID || activation == OAPH_FN_LINEAR ||
                 activation == OAPH_FN_TANH || activation == OAPH_FN_SIN,
             OAPH_ERR_ARG);
  Layer L;
  L.neurons = neurons;
  L.bias = has_bias ? 1u : 0u;
  L.activation = activation;
  const uint32_t t = L.total();
  OAPH_TRY(L.inputs.alloc(1, t));
  OAPH_TRY(L.sums.alloc(1, t));
  OAPH_TRY(L.errors.alloc(1, t));
  OAPH_TRY(L.errors_aux.alloc(1, t));
  OAPH_TRY(L.tinputs.alloc(t, 1));
  n->layers.push_back(L);
  Layer& cur = n->layers.back();
  OAPH_TRY(set_bias(cur));
  if (n->layers.size() >= 2) OAPH_TRY(connect(n->layers[n->layers.size() - 2], cur));
  if (n->layers.size() >= 2) {
    n->expected.release();
    OAPH_TRY(n->expected.alloc(1, cur.total()));
  }
  return OAPH_OK;
}

oaph_status oaph_network_set_learning_rate(oaph_network* n, double lr) {
  OAPH_CHECK(n, OAPH_ERR_ARG);
  n->learning_rate = lr;
  return OAPH_OK;
}

// weights of layer l -> l+1 as a math::Matrix: (total_l columns) x (neurons_{l+1}
// rows), element (c, r) at c + r * total_l.
oaph_status oaph_network_set_host_weights(oaph_network* n, uint32_t layer, const double* w) {
  OAPH_CHECK(n && w && layer + 1 < n->layers.size(), OAPH_ERR_ARG);
  OAPH_TRY(oaph_api_synchronize(n->api));
  return oaph_matrix_copy_h2d(n->layers[layer].weights.m, w, nullptr);
}

oaph_status oaph_network_get_host_weights(oaph_network* n, uint32_t layer, double* w) {
  OAPH_CHECK(n && w && layer + 1 < n->layers.size(), OAPH_ERR_ARG);
  OAPH_TRY(oaph_api_synchronize(n->api));
  return oaph_matrix_copy_d2h(w, nullptr, n->layers[layer].weights.m);
}

// setInputs (oapGenericNeuralApi.h:181-187): the first layer's neurons, bias 1
oaph_status oaph_network_set_inputs(oaph_network* n, const double* x) {
  OAPH_CHECK(n && x && n->layers.size() >= 2, OAPH_ERR_ARG);
  Layer& L = n->layers[0];
  n->host.assign(x, x + L.neurons);
  if (L.bias) n->host.push_back(1.0);
  OAPH_TRY(oaph_api_synchronize(n->api));
  return oaph_matrix_copy_h2d(L.inputs.m, n->host.data(), nullptr);
}

oaph_status oaph_network_set_expected(oaph_network* n, const double* t) {
  OAPH_CHECK(n && t && n->layers.size() >= 2, OAPH_ERR_ARG);
  const Layer& L = n->layers.back();
  n->host.assign(t, t + L.neurons);
  if (L.bias) n->host.push_back(0.0);
  OAPH_TRY(oaph_api_synchronize(n->api));
  return oaph_matrix_copy_h2d(n->expected.m, n->host.data(), nullptr);
}

// forwardPropagation (oapGenericNeuralApi.h:195-216)
oaph_status oaph_network_forward(oaph_network* n) {
  OAPH_CHECK(n && n->layers.size() >= 2, OAPH_ERR_ARG);
  for (size_t i = 1; i < n->layers.size(); ++i) {
    Layer& prev = n->layers[i - 1];
    Layer& cur = n->layers[i];
    const uint32_t dims[3][2] = {{1, cur.neurons}, {prev.total(), cur.neurons}, {1, prev.total()}};
    OAPH_TRY(oaph_dot_product_dims(n->api, cur.sums.m, prev.weights.m, prev.inputs.m, dims));
    OAPH_TRY(activation_dims(n->api, cur.activation, cur.inputs.m, cur.sums.m, 1, cur.neurons));
  }
  return OAPH_OK;
}

oaph_status oaph_network_get_outputs(oaph_network* n, double* y) {
  OAPH_CHECK(n && y && n->layers.size() >= 2, OAPH_ERR_ARG);
  const Layer& L = n->layers.back();
  n->host.resize(L.total());
  OAPH_TRY(oaph_api_synchronize(n->api));
  OAPH_TRY(oaph_matrix_copy_d2h(n->host.data(), nullptr, L.inputs.m));
  for (uint32_t i = 0; i < L.neurons; ++i) y[i] = n->host[i];
  return OAPH_OK;
}

// accumulateErrors with MEAN_SQUARE_ERROR (oapNetwork.cpp:335-352,
// oapGenericNeuralApi.h:245-262): e = y - t on the device, 0.5 e^2 kept per
// output for calculateError.
oaph_status oaph_network_accumulate_errors(oaph_network* n) {
  OAPH_CHECK(n && n->layers.size() >= 2, OAPH_ERR_ARG);
  Layer& L = n->layers.back();
  OAPH_TRY(oaph_substract(n->api, L.errors_aux.m, L.inputs.m, n->expected.m));
  n->host.resize(L.total());
  OAPH_TRY(oaph_api_synchronize(n->api));
  OAPH_TRY(oaph_matrix_copy_d2h(n->host.data(), nullptr, L.errors_aux.m));
  for (uint32_t i = 0; i < L.total(); ++i) n->errors.push_back(n->host[i] * n->host[i] * 0.5);
  return OAPH_OK;
}

// backPropagation (oapGenericNeuralApi.h:264-334)
oaph_status oaph_network_backward(oaph_network* n) {
  OAPH_CHECK(n && n->layers.size() >= 2, OAPH_ERR_ARG);
  oaph_api* api = n->api;
  auto layer_errors = [&](Layer& cur) -> oaph_status {
    OAPH_TRY(activation_dims(api, derivative_of(cur.activation), cur.sums.m, cur.sums.m, 1,
                             cur.neurons));
    return oaph_hadamard_product_vec(api, cur.errors.m, cur.errors.m, cur.sums.m);
  };
  const size_t last = n->layers.size() - 1;
  {
    const oaph_matrix* src = n->layers[last].errors_aux.m;
    OAPH_HIP(hipMemcpyAsync(n->layers[last].errors.m->reValues, src->reValues,
                            numel(src) * sizeof(double), hipMemcpyDeviceToDevice, api->stream));
  }
  OAPH_TRY(layer_errors(n->layers[last]));
  for (size_t idx = last; idx > 1; --idx) {
    Layer& next = n->layers[idx];
    Layer& cur = n->layers[idx - 1];
    OAPH_TRY(oaph_transpose(api, cur.tweights.m, cur.weights.m));
    const uint32_t dims[3][2] = {{1, cur.total()}, {next.neurons, cur.total()}, {1, next.neurons}};
    OAPH_TRY(oaph_dot_product_dims(api, cur.errors.m, cur.tweights.m, next.errors.m, dims));
    OAPH_TRY(layer_errors(cur));
  }
  for (size_t idx = 0; idx + 1 < n->layers.size(); ++idx) {
    Layer& cur = n->layers[idx];
    Layer& next = n->layers[idx + 1];
    OAPH_TRY(oaph_transpose(api, cur.tinputs.m, cur.inputs.m));
    const uint32_t dims[3][2] = {{cur.total(), next.neurons}, {cur.total(), 1}, {1, next.neurons}};
    OAPH_TRY(oaph_tensor_product(api, cur.weights1.m, cur.tinputs.m, next.errors.m, dims));
    OAPH_TRY(oaph_add(api, cur.weights2.m, cur.weights2.m, cur.weights1.m));
  }
  return OAPH_OK;
}

// postStep (oapNetwork.cpp:553-567): nabla sums and the error list cleared
oaph_status oaph_network_post_step(oaph_network* n) {
  OAPH_CHECK(n, OAPH_ERR_ARG);
  for (size_t idx = 0; idx + 1 < n->layers.size(); ++idx)
    OAPH_TRY(oaph_set_zero_matrix(n->api, n->layers[idx].weights2.m));
  n->errors.clear();
  return OAPH_OK;
}

// updateWeights (oapGenericNeuralApi.h:336-354), then postStep
oaph_status oaph_network_update_weights(oaph_network* n) {
  OAPH_CHECK(n && n->layers.size() >= 2, OAPH_ERR_ARG);
  OAPH_CHECK(!n->errors.empty(), OAPH_ERR_STATE);
  const double lr = n->learning_rate / (double)n->errors.size();
  for (size_t idx = 0; idx + 1 < n->layers.size(); ++idx) {
    Layer& cur = n->layers[idx];
    OAPH_TRY(oaph_multiply_re_constant(n->api, cur.weights2.m, cur.weights2.m, lr));
    OAPH_TRY(oaph_substract(n->api, cur.weights.m, cur.weights.m, cur.weights2.m));
  }
  return oaph_network_post_step(n);
}

// calculateError(MEAN_SQUARE_ERROR) (oapNetwork.cpp:280-285): mean of 0.5 e^2
oaph_status oaph_network_calculate_error(oaph_network* n, double* out) {
  OAPH_CHECK(n && out, OAPH_ERR_ARG);
  OAPH_CHECK(!n->errors.empty(), OAPH_ERR_STATE);
  double s = 0;
  for (double e : n->errors) s += e;
  *out = s / (double)n->errors.size();
  return OAPH_OK;
}

}  // extern "C"
