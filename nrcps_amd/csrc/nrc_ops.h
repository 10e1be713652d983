// nrc_ops.h -- per-site multipliers and the plane-wave sink gather, shared by the pass
// kernels (fft_passes.hip) and the single-pass slice kernel (slice_fused.hip).
#pragma once

#include "nrc_device.h"
#include "nrc_internal.h"

namespace nrc {

// ------------------------------------------------------------------ per-site multipliers
// v * (hi + lo): hi carries the factor, lo its FP32 remainder (zero in FP64 passes)
template <typename T>
__device__ __forceinline__ C2<T> scale2(C2<T> v, T hi, T lo) {
  if constexpr (sizeof(T) == 4) {
    return {fmaf(v.x, lo, v.x * hi), fmaf(v.y, lo, v.y * hi)};
  } else {
    return {v.x * hi, v.y * hi};
  }
}

// v[i] *= multiplier at site(i) = i * stride + off, the kind fixed at compile time
template <typename T, int L, int KIND>
__device__ __forceinline__ void apply_mul_k(C2<T>* v, const MulSpec& m, long cfg, long off,
                                            long stride, long V) {
  if constexpr (KIND == MUL_FIELD) {
    const double dp = m.a0 + m.a1, dm = m.a0 - m.a1;
    const T ph = T(dp), mh = T(dm);
    const T pl = T(dp - (double)ph), ml = T(dm - (double)mh);
    const uint32_t* bits = m.bits + cfg * m.nw;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const long site = i * stride + off;
      const bool up = (bits[site >> 5] >> (site & 31)) & 1u;
      v[i] = scale2<T>(v[i], up ? ph : mh, up ? pl : ml);
    }
  } else if constexpr (KIND == MUL_TABLE || KIND == MUL_RFIELD) {
    const T* tab = reinterpret_cast<const T*>(m.table) + (KIND == MUL_RFIELD ? cfg * V : 0);
    const T* lo = reinterpret_cast<const T*>(m.table_lo);
    // chunks of 8 sites with scheduling barriers between them: the table loads of later
    // chunks are not hoisted, so at most 16 table registers are live at a time
#pragma unroll
    for (int i0 = 0; i0 < L; i0 += 8) {
#pragma unroll
      for (int i = i0; i < i0 + 8 && i < L; ++i) {
        const long site = i * stride + off;
        v[i] = scale2<T>(v[i], tab[site], (KIND == MUL_TABLE && lo) ? lo[site] : T(0));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// v[i] *= multiplier at site(i) = i * stride + off
template <typename T, int L>
__device__ __forceinline__ void apply_mul(C2<T>* v, const MulSpec& m, long cfg, long off,
                                          long stride, long V) {
  if (m.kind == MUL_NONE) return;
  if (m.kind == MUL_FIELD) {
    const double dp = m.a0 + m.a1, dm = m.a0 - m.a1;
    const T ph = T(dp), mh = T(dm);
    const T pl = T(dp - (double)ph), ml = T(dm - (double)mh);
    const uint32_t* bits = m.bits + cfg * m.nw;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const long site = i * stride + off;
      const bool up = (bits[site >> 5] >> (site & 31)) & 1u;
      v[i] = scale2<T>(v[i], up ? ph : mh, up ? pl : ml);
    }
    return;
  }
  const T* tab = reinterpret_cast<const T*>(m.table) + (m.kind == MUL_RFIELD ? cfg * V : 0);
  const T* lo = reinterpret_cast<const T*>(m.table_lo);
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const long site = i * stride + off;
    v[i] = scale2<T>(v[i], tab[site], lo ? lo[site] : T(0));
  }
}

template <typename T, int L>
__device__ __forceinline__ void gather_pencil(const C2<T>* v, const GatherSpec& g, int pencil,
                                              long cfg, int orb, int norb) {
  const int e0 = g.ptr[pencil], e1 = g.ptr[pencil + 1];
  for (int e = e0; e < e1; ++e) {
    const C2<T> val = reg_select<T, L>(v, g.pos[e]);
    g.M[(cfg * g.nsink + g.sink[e]) * norb + orb] =
        make_double2((double)val.x * g.scale, (double)val.y * g.scale);
  }
}

}  // namespace nrc
