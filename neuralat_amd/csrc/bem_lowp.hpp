// Launchers of the low-precision / matrix-core kernels (bem_lowp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bem {

constexpr int kDtBF16 = 1;
constexpr int kDtFP16 = 2;
constexpr int kDtFP8 = 3;  // OCP e4m3 with per-row scales (batched mixed-precision solve only)
constexpr int kMaxSplit = 8;  // column splits of the multi-RHS products

hipError_t launch_absmax(const float2* A, int64_t rows, int64_t cols, int64_t lda, float* partial, int nblocks,
                         hipStream_t s);
hipError_t launch_to_lowp(int dt, const float2* A, int64_t rows, int64_t cols, int64_t lda, float scale, uint16_t* Ar,
                          uint16_t* Ai, int64_t ldo, hipStream_t s);
hipError_t launch_vec_to_lowp(int dt, const float2* X, int64_t n, int64_t ldx_in, int nrhs, uint16_t* Xr, uint16_t* Xi,
                              int64_t ldx, hipStream_t s);
hipError_t launch_gemm_lowp(int dt, const uint16_t* Ar, const uint16_t* Ai, int64_t rows, int64_t ld, float inv_scale,
                            const uint16_t* Xr, const uint16_t* Xi, int64_t ldx, int nrhs, float2* Y, int64_t ldy,
                            float2 alpha, float2 beta, float2* part, hipStream_t s);
// part: scratch of splitk_part_elems(rows) complex values for split-K partial
// sums (deterministic fixed-order reduction), or null for a single pass.
hipError_t launch_gemm_c64(const float2* A, int64_t rows, int64_t cols, int64_t lda, const float2* X, int64_t ldx,
                           int nrhs, float2* Y, int64_t ldy, float2 alpha, float2 beta, float2* part, hipStream_t s);
int64_t splitk_part_elems(int64_t rows);
// entries of one tiled 16-bit plane of a rows x cols matrix with column pitch ld
int64_t lowp_plane_elems(int64_t rows, int64_t ld);

}  // namespace bem
