// Learned coarse correction (4-level 3D U-Net on the level-1 node grid): internal interface
// between the network kernels and the solver.
#pragma once
#include <functional>

#include "dfmi_internal.h"

namespace dfmi {

struct NetDims {
  int x, y, z;
  __host__ __device__ long long n() const { return (long long)x * y * z; }
};

struct NetImpl;
NetImpl* net_new();
void net_delete(NetImpl* N);
void net_set_stream(NetImpl* N, hipStream_t s);
long long net_param_count(const NetImpl* N);
void net_set_params(NetImpl* N, const float* src);       // FP32 master weights (device), resets Adam
void net_get(NetImpl* N, int which, float* dst);          // 0: master weights, 1: gradients
void net_forward(NetImpl* N, int n, const NetDims& d1, const float* x);  // x (n, 6, d1) -> output (n, 3, d1)
const float* net_output(NetImpl* N);
void net_copy_output(NetImpl* N, float* dst, long long count);  // synchronises the network stream
void net_backward(NetImpl* N, const float* dout);         // gradients of the last forward
void net_adam(NetImpl* N, double lr, double b1, double b2, double eps);
void net_invalidate_feature(NetImpl* N);
// delta = Dinv g + Dinv^(1/2) s N(Dinv^(1/2) g / s, f) on n level-1 vectors in layout lay
void net_correct(NetImpl* N, const Geom& g1, const float* g, VLay lay, int n, const float* dinv, const int* active,
                 float* delta);
// energy-optimal step per vector: delta *= max(0, delta.g / delta.A delta)
void net_energy_step(NetImpl* N, const float* g, const float* ad, float* delta, VLay lay, long long nodes, int n,
                     const int* active);
// one Adam step on the normalised energy loss; returns the summed loss
double net_train_step(NetImpl* N, const Geom& g1, const float* g, int n, const float* dinv, double lr,
                      const std::function<void(const float*, float*)>& apply);
void gather_samples(const float* src, VLay lay, long long nodes, int n, const int* active, float* dst,
                    hipStream_t s);

}  // namespace dfmi
