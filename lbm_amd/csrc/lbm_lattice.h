// Lattice tables and storage traits shared by the HIP kernels.
//
// Velocity sets, weights and opposite maps follow the reference's LBMethod
// specialisations (src/lbm/constants.h:295-319 for D2Q9 and :321-365 for
// D3Q19): the rest population is the last index and the D3Q19 xy-diagonals
// are ordered differently from D2Q9.  The tables are constexpr so that the
// fully unrolled population loops fold every c_i and w_i into immediates.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

template <int Q>
struct Lat;

template <>
struct Lat<9> {
  static constexpr int D = 2;
  static constexpr int cx[9] = {-1, 1, 0, 0, 1, 1, -1, -1, 0};
  static constexpr int cy[9] = {0, 0, -1, 1, 1, -1, -1, 1, 0};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {1, 0, 3, 2, 6, 7, 4, 5, 8};
  static constexpr double w[9] = {1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 4.0 / 9.0};
};

template <>
struct Lat<19> {
  static constexpr int D = 3;
  static constexpr int cx[19] = {-1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0};
  static constexpr int cz[19] = {0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, 0};
  static constexpr int opp[19] = {1, 0, 3, 2, 5, 4, 9, 8, 7, 6, 13, 12, 11, 10, 17, 16, 15, 14, 18};
  static constexpr double w[19] = {1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                   1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0};
};

// Storage traits: populations are stored shifted, h_i = f_i - w_i.
//   R: the type populations are held in registers (exact copy of storage
//      for 32/64-bit storage, FP32 for 16-bit storage);
//   C: the arithmetic type of moments, equilibrium and collision.
// FP32 storage computes in FP64: the non-equilibrium part heq - h is a small
// difference of two O(w) numbers, and in FP32 its rounding error is
// correlated from step to step in smooth flows (2e-5 relative velocity error
// after 1000 steps of the 256^3 Taylor-Green box, against 2e-8 with FP64
// arithmetic on the same FP32 storage).  16-bit storage computes in FP32.
template <typename ST>
struct Store;

#ifndef LBM_F32_COMPUTE_F64
#define LBM_F32_COMPUTE_F64 1
#endif
template <>
struct Store<float> {
  using R = float;
#if LBM_F32_COMPUTE_F64
  using C = double;
#else
  using C = float;
#endif
  static __device__ __forceinline__ R ld(const float *p) { return *p; }
  static __device__ __forceinline__ R ld_nt(const float *p) { return __builtin_nontemporal_load(p); }
  static __device__ __forceinline__ void st(float *p, C v) { *p = float(v); }
  static __device__ __forceinline__ void st_nt(float *p, C v) { __builtin_nontemporal_store(float(v), p); }
};

template <>
struct Store<double> {
  using R = double;
  using C = double;
  static __device__ __forceinline__ C ld(const double *p) { return *p; }
  static __device__ __forceinline__ C ld_nt(const double *p) { return __builtin_nontemporal_load(p); }
  static __device__ __forceinline__ void st(double *p, C v) { *p = v; }
  static __device__ __forceinline__ void st_nt(double *p, C v) { __builtin_nontemporal_store(v, p); }
};

// 16-bit forms move the raw bits as uint16_t so that plain (and nontemporal)
// 2-byte global accesses are used; conversion is round-to-nearest-even.
template <>
struct Store<__half> {
  using R = float;
  using C = float;
  static __device__ __forceinline__ C ld(const __half *p) {
    return __half2float(__ushort_as_half(*reinterpret_cast<const unsigned short *>(p)));
  }
  static __device__ __forceinline__ C ld_nt(const __half *p) {
    return __half2float(__ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p))));
  }
  static __device__ __forceinline__ void st(__half *p, C v) {
    *reinterpret_cast<unsigned short *>(p) = __half_as_ushort(__float2half_rn(v));
  }
  static __device__ __forceinline__ void st_nt(__half *p, C v) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v)), reinterpret_cast<unsigned short *>(p));
  }
};

template <>
struct Store<__hip_bfloat16> {
  using R = float;
  using C = float;
  static __device__ __forceinline__ float bits2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }
  static __device__ __forceinline__ unsigned short f2bits(float v) {
    // round to nearest even on the upper 16 bits; NaN kept quiet
    unsigned int u = __float_as_uint(v);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
  }
  static __device__ __forceinline__ C ld(const __hip_bfloat16 *p) {
    return bits2f(*reinterpret_cast<const unsigned short *>(p));
  }
  static __device__ __forceinline__ C ld_nt(const __hip_bfloat16 *p) {
    return bits2f(__builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p)));
  }
  static __device__ __forceinline__ void st(__hip_bfloat16 *p, C v) {
    *reinterpret_cast<unsigned short *>(p) = f2bits(v);
  }
  static __device__ __forceinline__ void st_nt(__hip_bfloat16 *p, C v) {
    __builtin_nontemporal_store(f2bits(v), reinterpret_cast<unsigned short *>(p));
  }
};

}  // namespace lbm
