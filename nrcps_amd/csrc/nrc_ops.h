// nrc_ops.h -- per-site multipliers and the plane-wave sink gather, shared by the pass
// kernels (fft_passes.hip) and the single-pass slice kernel (slice_fused.hip).
#pragma once

#include "nrc_device.h"
#include "nrc_internal.h"

namespace nrc {

// ------------------------------------------------------------------ per-site multipliers
// v * (hi + lo): hi carries the factor, lo its FP32 remainder (zero in FP64 passes); the
// remainder product enters the fma of the main product, before its one rounding
template <typename T>
__device__ __forceinline__ C2<T> scale2(C2<T> v, T hi, T lo) {
  if constexpr (sizeof(T) == 4) {
    return {fmaf(v.x, hi, v.x * lo), fmaf(v.y, hi, v.y * lo)};
  } else {
    return {v.x * hi, v.y * hi};
  }
}

// A MUL_TABLE table holds one TabEntry per site: (hi, lo) float pairs in FP32 passes, the
// double itself in FP64 passes.
template <typename T>
struct TabEntryT {
  using type = float2;
};
template <>
struct TabEntryT<double> {
  using type = double;
};
template <typename T>
using TabEntry = typename TabEntryT<T>::type;

template <typename T>
__device__ __forceinline__ C2<T> scale_tab(C2<T> v, TabEntry<T> e) {
  if constexpr (sizeof(T) == 4) {
    return scale2<T>(v, e.x, e.y);
  } else {
    return scale2<T>(v, e, T(0));
  }
}

// ------------------------------------------------------------------ buffer addressing
// Global accesses of the slice kernels go through buffer resources: the uniform base lives in
// SGPRs, each lane adds one 32-bit byte offset and each access a constant byte offset held in
// an SGPR, so a whole pencil of loads needs a single address VGPR.  Accesses past `bytes`
// read zero and are dropped on store.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kBufWord3 = 0x00020000;  // raw 32-bit accesses (gfx9 buffer descriptor word 3)

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kBufWord3);
}

template <typename U>
__device__ __forceinline__ U bload(rsrc_t r, uint32_t voff, uint32_t soff) {
  static_assert(sizeof(U) == 4 || sizeof(U) == 8 || sizeof(U) == 16, "buffer load size");
  if constexpr (sizeof(U) == 4) {
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else if constexpr (sizeof(U) == 8) {
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
}

template <typename U>
__device__ __forceinline__ void bstore(U val, rsrc_t r, uint32_t voff, uint32_t soff) {
  static_assert(sizeof(U) == 8 || sizeof(U) == 16, "buffer store size");
  if constexpr (sizeof(U) == 8) {
    using V2 = __attribute__((__vector_size__(2 * sizeof(uint32_t)))) uint32_t;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, val), r, voff, soff, 0);
  } else {
    using V4 = __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, val), r, voff, soff, 0);
  }
}

// v[i] *= multiplier at site(i) = i * STRIDE + off for i < L, the kind fixed at compile time.
// Sites of one call differ by the compile-time STRIDE only, so table and field reads take one
// lane offset plus constants.  V is the volume of one configuration.
template <typename T, int L, long STRIDE, int KIND>
__device__ __forceinline__ void apply_mul_k(C2<T>* v, const MulSpec& m, long cfg, uint32_t off,
                                            long V) {
  if constexpr (KIND == MUL_FIELD) {
    const double dp = m.a0 + m.a1, dm = m.a0 - m.a1;
    const T ph = T(dp), mh = T(dm);
    const T pl = T(dp - (double)ph), ml = T(dm - (double)mh);
    const rsrc_t rb = make_rsrc(m.bits + cfg * m.nw, (uint32_t)m.nw * 4u);
    if (m.table) {  // with a trap: the factors (a0 + a1 phi) q(site), phi = +1 table first
      using E = TabEntry<T>;
      const rsrc_t rt = make_rsrc(m.table, (uint32_t)(2 * V * sizeof(E)));
#pragma unroll
      for (int i = 0; i < L; ++i) {
        const uint32_t site = (uint32_t)(i * STRIDE) + off;
        const uint32_t w = bload<uint32_t>(rb, (site >> 5) << 2, 0u);
        const bool up = (w >> (site & 31u)) & 1u;
        const uint32_t idx = site + (up ? 0u : (uint32_t)V);
        v[i] = scale_tab<T>(v[i], bload<E>(rt, idx * (uint32_t)sizeof(E), 0u));
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
      uint32_t w, bit;
      if constexpr (STRIDE % 32 == 0) {  // word = i * STRIDE / 32 + off / 32
        w = bload<uint32_t>(rb, (off >> 5) << 2, (uint32_t)(i * (STRIDE / 32)) * 4u);
        bit = off & 31u;
      } else {
        const uint32_t site = (uint32_t)(i * STRIDE) + off;
        w = bload<uint32_t>(rb, (site >> 5) << 2, 0u);
        bit = site & 31u;
      }
      const bool up = (w >> bit) & 1u;
      v[i] = scale2<T>(v[i], up ? ph : mh, up ? pl : ml);
    }
  } else if constexpr (KIND == MUL_TABLE) {
    using E = TabEntry<T>;
    const rsrc_t rt = make_rsrc(m.table, (uint32_t)(V * sizeof(E)));
    // chunks of 8 sites with scheduling barriers between them: the table loads of later
    // chunks are not hoisted, so few table registers are live at a time
#pragma unroll
    for (int i0 = 0; i0 < L; i0 += 8) {
#pragma unroll
      for (int i = i0; i < i0 + 8 && i < L; ++i)
        v[i] = scale_tab<T>(v[i], bload<E>(rt, off * (uint32_t)sizeof(E),
                                           (uint32_t)(i * STRIDE * (long)sizeof(E))));
      __builtin_amdgcn_sched_barrier(0);
    }
  } else if constexpr (KIND == MUL_RFIELD) {
    const rsrc_t rt = make_rsrc(reinterpret_cast<const T*>(m.table) + cfg * V,
                                (uint32_t)(V * sizeof(T)));
#pragma unroll
    for (int i0 = 0; i0 < L; i0 += 8) {
#pragma unroll
      for (int i = i0; i < i0 + 8 && i < L; ++i) {
        const T f = bload<T>(rt, off * (uint32_t)sizeof(T), (uint32_t)(i * STRIDE * (long)sizeof(T)));
        v[i] = scale2<T>(v[i], f, T(0));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// v[i] *= multiplier at site(i) = i * stride + off (the kind chosen at run time)
template <typename T, int L>
__device__ __forceinline__ void apply_mul(C2<T>* v, const MulSpec& m, long cfg, long off,
                                          long stride, long V) {
  if (m.kind == MUL_NONE) return;
  if (m.kind == MUL_FIELD) {
    const double dp = m.a0 + m.a1, dm = m.a0 - m.a1;
    const T ph = T(dp), mh = T(dm);
    const T pl = T(dp - (double)ph), ml = T(dm - (double)mh);
    const uint32_t* bits = m.bits + cfg * m.nw;
    const TabEntry<T>* trap = reinterpret_cast<const TabEntry<T>*>(m.table);
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const long site = i * stride + off;
      const bool up = (bits[site >> 5] >> (site & 31)) & 1u;
      if (trap)  // (a0 + a1 phi) q(site), phi = +1 table first
        v[i] = scale_tab<T>(v[i], trap[site + (up ? 0 : V)]);
      else
        v[i] = scale2<T>(v[i], up ? ph : mh, up ? pl : ml);
    }
    return;
  }
  if (m.kind == MUL_RFIELD) {
    const T* tab = reinterpret_cast<const T*>(m.table) + cfg * V;
#pragma unroll
    for (int i = 0; i < L; ++i) v[i] = scale2<T>(v[i], tab[i * stride + off], T(0));
    return;
  }
  const TabEntry<T>* tab = reinterpret_cast<const TabEntry<T>*>(m.table);
#pragma unroll
  for (int i = 0; i < L; ++i) v[i] = scale_tab<T>(v[i], tab[i * stride + off]);
}

// w[y] *= (a0 + a1 phi) q(site), site = off + y L, phi from the field words in LDS, with a trap: w[y] *= (a0 + a1 phi) q(site) read from the two factor tables
// (phi = +1 first, MUL_TABLE format), four sites at a time so that few loads are in flight
template <typename T, int L>
__device__ __forceinline__ void mul_field_trap(C2<T>* w, const uint32_t* bl, uint32_t off,
                                               rsrc_t rt, uint32_t V) {
  using E = TabEntry<T>;
  constexpr int G = 4;
#pragma unroll
  for (int y0 = 0; y0 < L; y0 += G) {
    E e[G];
#pragma unroll
    for (int y = y0; y < y0 + G && y < L; ++y) {
      const uint32_t site = off + (uint32_t)(y * L);
      const bool up = (bl[site >> 5] >> (site & 31u)) & 1u;
      e[y - y0] = bload<E>(rt, (site + (up ? 0u : V)) * (uint32_t)sizeof(E), 0u);
    }
#pragma unroll
    for (int y = y0; y < y0 + G && y < L; ++y) w[y] = scale_tab<T>(w[y], e[y - y0]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename T, int L>
__device__ __forceinline__ void gather_pencil(const C2<T>* v, const GatherSpec& g, int pencil,
                                              long cfg, int orb, int norb) {
  const int nc = g.col ? g.ncol : norb;
  const int c0 = g.col ? g.col[orb] : orb;
  const int e0 = g.ptr[pencil], e1 = g.ptr[pencil + 1];
  for (int e = e0; e < e1; ++e) {
    const C2<T> val = reg_select<T, L>(v, g.pos[e]);
    const double f = g.sink_scale ? g.scale * g.sink_scale[g.sink[e]] : g.scale;
    g.M[(cfg * g.nsink + g.sink[e]) * nc + c0] =
        make_double2((double)val.x * f, (double)val.y * f);
  }
  if (!g.col) return;
  const int pc = g.pcol[orb];
  if (pc < 0) return;
  const int m0 = g.ptr2[pencil], m1 = g.ptr2[pencil + 1];
  for (int e = m0; e < m1; ++e) {
    const C2<T> val = reg_select<T, L>(v, g.pos2[e]);
    const double f = g.sink_scale ? g.scale * g.sink_scale[g.sink2[e]] : g.scale;
    g.M[(cfg * g.nsink + g.sink2[e]) * nc + pc] =
        make_double2((double)val.x * f, -(double)val.y * f);
  }
}

}  // namespace nrc
