// bfa_mlp.h -- internal interface of the surrogate MLP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define MLP_KSPLIT 16

namespace bfa {

struct MLPShape {
  int nin, kp, h, nout;
  int64_t oW[4], ob[4];   // offsets of the weight matrices / biases in a member's parameter block
  int64_t nparam;
};

MLPShape mlp_shape(int nin, int hidden, int nout);
int64_t mlp_work_floats(const MLPShape &s, int B);   // per member
int64_t mlp_work_halfs(const MLPShape &s, int B);    // per member
// X16 [E][B][kp] (member stride sx; 0 = shared), W16 [E][nparam], bias32 = FP32 parameter block [E][nparam]
int mlp_forward(const MLPShape &s, int E, int B, const _Float16 *X16, int64_t sx, const _Float16 *W16,
                const float *bias32, float *Y, float *Zsave, _Float16 *Abuf, hipStream_t st);
int mlp_backward(const MLPShape &s, int E, int B, const _Float16 *X16, int64_t sx, const _Float16 *W16,
                 const float *Zsave, const _Float16 *Abuf, const float *dY, float *grad, float *work, hipStream_t st);
// w: per-sample weights [E][B] or NULL
int mlp_mse(int E, int B, int nout, const float *Y, const float *T, const float *w, float *dY, double *loss_parts,
            int nparts, hipStream_t st);
int mlp_adam(int64_t n, float *w, const float *g, float *m, float *v, _Float16 *w16, float lr, int step,
             hipStream_t st);
int to_half(const float *src, _Float16 *dst, int64_t n, hipStream_t st);

}  // namespace bfa
