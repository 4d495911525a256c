// Learned shortlist: the 128 -> 256 -> 32 embedding network and its code pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace v3dm {

constexpr int kHid = 256;
constexpr int kCode = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * 128;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kCode * kHid;
constexpr int kParams = kOffB2 + kCode;  // 41248
constexpr int kRowPadMlp = 512;          // minibatch / embedding rows are padded to this

// Device parameters: FP32 masters (P, the packed vector) and the bf16 MFMA copies.
struct MlpWeights {
    float* P = nullptr;
    __bf16* W1b = nullptr;   // [256][128]
    __bf16* W2b = nullptr;   // [32][256]
    __bf16* W2Tb = nullptr;  // [256][32]
    __bf16* W2p = nullptr;   // W2 as the A operands of the inference kernel's second layer
};

void launch_mlp_gather(float const* pool, int const* tri, int const* order, int orderOff, int B, int Rp,
                       __bf16* X, __bf16* XT, hipStream_t st);
void launch_mlp_rows_bf16(float const* src, int n, int Rp, __bf16* X, hipStream_t st);
void launch_mlp_forward(__bf16 const* X, int Rp, MlpWeights const& w, __bf16* HT, float* Z, hipStream_t st);
// Inference: quantised codes of FP32 rows [0, rows128) (a multiple of 128 rows, all
// readable), 32 bytes per row in the kernel's own (fixed) order of the code dimensions.
void launch_mlp_codes_f32(float const* X, int rows128, MlpWeights const& w, float scale, signed char* codes,
                          hipStream_t st);
// uint8 rows with their dequantisation scales (element = (float)q * denom[row]): the same
// codes as the FP32 rows they dequantise to
void launch_mlp_codes_u8(uint8_t const* Q, float const* denom, int rows128, MlpWeights const& w, float scale,
                         signed char* codes, hipStream_t st);
// two feature sets in one launch (row counts multiples of 128; either may be 0)
void launch_mlp_codes2_f32(float const* X0, int rows0, signed char* codes0, float const* X1, int rows1,
                           signed char* codes1, MlpWeights const& w, float scale, hipStream_t st);
void launch_mlp_loss(float const* Z, int B, float tau, float* dC, int* hard, float* lossPart, hipStream_t st);
void launch_mlp_norm_backward(float const* Z, float const* dC, int R, int Rp, float* dZf, __bf16* dZ, __bf16* dZT,
                              hipStream_t st);
void launch_mlp_backward_h(__bf16 const* dZ, int Rp, MlpWeights const& w, __bf16 const* HT, __bf16* dHT,
                           hipStream_t st);
int mlp_wgrad_chunks(int Rp);
void launch_mlp_wgrad(__bf16 const* XT, __bf16 const* HT, __bf16 const* dZT, __bf16 const* dHT, float const* dZf,
                      int Rp, float* part, hipStream_t st);
// apply = 0: only sums the gradient (and refreshes the bf16 copies from the masters)
void launch_mlp_adam(float const* part, int nChunks, float* grad, MlpWeights const& w, float* M, float* V, float lr,
                     int step, int apply, hipStream_t st);
void launch_fill(float* p, size_t n, float v, hipStream_t st);

}  // namespace v3dm
