// Internal definition of the MLP handle (mlp.hip) shared with the learned reactor step.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

struct flint_mlp {
  int L = 0;                               // layers
  std::vector<int> dims;                   // widths d_0..d_L
  std::vector<long> w_off, b_off, w16_off; // offsets of W_l, b_l in the FP32 vector, of W_l in the BF16 copy
  long np = 0, nw = 0;                     // parameters, weights
  float *w32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;  // master, gradient, Adam moments
  __bf16* w16 = nullptr;                   // BF16 weights
  __bf16* w16t = nullptr;                  // BF16 weights transposed per layer ([dout][din]) for the forward path
  std::vector<long> w16t_off;
  bool wt_stale = true;                    // w16t behind w16 (set on every weight update)
  unsigned long long uid = 0;              // unique over the process (flint_mlp_create)
  unsigned long long version = 0;          // counts weight updates (copies derived elsewhere check both)
  int* w16_index = nullptr;                // FP32 index -> BF16 index (-1 for biases)
  float* loss = nullptr;                   // device scalar
  int adam_t = 0;
  // batch workspace
  long cap = 0;
  bool cap_train = false;
  std::vector<void*> ws;
  std::vector<__bf16*> a16;                // input of layer l (a16[0] = network input)
  std::vector<float*> z32;                 // pre-activation of layer l
  std::vector<float*> dz32;
  std::vector<__bf16*> dz16;

  ~flint_mlp();
  int reserve(long n, bool train);
  int load_input(long n, const float* X, hipStream_t s);
  int refresh_wt(hipStream_t s);
  int forward(long n, float* out, hipStream_t s, bool keep = true);
  int backward(long n, const float* t, hipStream_t s);
};
