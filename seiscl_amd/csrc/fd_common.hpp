// Device helpers shared by the finite-difference kernels.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "seis_internal.hpp"

namespace seis {
namespace {

constexpr int BZ = 64;   // lanes along z
constexpr int BT = 4;    // rows along y (3D) or x (2D)

// Forward / backward staggered derivatives, src/header_FD.cl:87-265.
template <int H, typename R>
__device__ __forceinline__ R Dp(const R *__restrict__ v, long i, long s, const R *hc) {
    R acc = hc[1] * (v[i + s] - v[i]);
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (v[i + k * s] - v[i - (k - 1) * s]);
    return acc;
}
template <int H, typename R>
__device__ __forceinline__ R Dm(const R *__restrict__ v, long i, long s, const R *hc) {
    R acc = hc[1] * (v[i] - v[i - s]);
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (v[i + (k - 1) * s] - v[i - k * s]);
    return acc;
}

// CPML coefficient access: cpml[axis][kind][2*nab]
enum { CK = 0, CA = 1, CB = 2, CKH = 3, CAH = 4, CBH = 5 };
template <typename R>
__device__ __forceinline__ R cp(const FDArgs<R> &a, int axis, int kind, int i) {
    return a.cpml[(axis * 6 + kind) * 2 * a.nab + i];
}

// psi <- b psi + a d ;  d <- d / K + psi   (src/update_v3D.cl:210-314)
template <typename R>
__device__ __forceinline__ void pml(R &d, R *psi, R b, R aa, R K) {
    R ps = b * *psi + aa * d;
    *psi = ps;
    d = d / K + ps;
}

// CPML memory update split in two phases so that every load of a cell is
// issued before its first store: pml_load fetches psi and the three
// profile coefficients, pml_apply does psi <- b psi + a d, d <- d/K + psi.
template <typename R>
struct Pml {
    R psi, a, b, K;
};
// NT: non-temporal access to the memory variable.
template <bool NT = false, typename R>
__device__ __forceinline__ Pml<R> pml_load(const FDArgs<R> &a, const R *psi, int axis, bool half,
                                           int idx) {
    Pml<R> p;
    const int k0 = half ? CKH : CK;
    if constexpr (NT) p.psi = __builtin_nontemporal_load(psi);
    else p.psi = *psi;
    p.K = cp(a, axis, k0, idx);
    p.a = cp(a, axis, k0 + 1, idx);
    p.b = cp(a, axis, k0 + 2, idx);
    return p;
}
template <bool NT = false, typename R>
__device__ __forceinline__ void pml_apply(R &d, R *psi, const Pml<R> &p) {
    const R ps = p.b * p.psi + p.a * d;
    if constexpr (NT) __builtin_nontemporal_store(ps, psi);
    else *psi = ps;
    d = d / p.K + ps;
}

// Renumber blocks: hardware block b runs on XCD b % 8; give each XCD a
// contiguous slab of tiles.
__device__ __forceinline__ void xcd_tile(unsigned nb, unsigned &tile) {
    const unsigned b = blockIdx.x;
    const unsigned per = (nb + 7) / 8;
    const unsigned xcd = b % 8, slot = b / 8;
    tile = xcd * per + slot;
}


// x-derivatives from a register window w[j] = f[x - H + j], j = 0..2H,
// with the same terms and order as Dp / Dm.
template <int H, typename R>
__device__ __forceinline__ R Dpw(const R (&w)[2 * H + 1], const R *hc) {
    R acc = hc[1] * (w[H + 1] - w[H]);
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (w[H + k] - w[H - k + 1]);
    return acc;
}
template <int H, typename R>
__device__ __forceinline__ R Dmw(const R (&w)[2 * H + 1], const R *hc) {
    R acc = hc[1] * (w[H] - w[H - 1]);
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (w[H + k - 1] - w[H - k]);
    return acc;
}
template <int H, typename R>
__device__ __forceinline__ void wshift(R (&w)[2 * H + 1]) {
#pragma unroll
    for (int j = 0; j < 2 * H; ++j) w[j] = w[j + 1];
}

// Loads and stores addressed as (uniform base pointer) + (32-bit per-lane
// byte offset): every neighbour of a cell shares the lane's single offset
// register and differs only in the scalar base, which keeps the marching
// kernels' VGPR budget low.  Each array is below 4 GiB (checked on the host).
template <typename R>
__device__ __forceinline__ R ldo(const R *base, unsigned boff) {
    return *reinterpret_cast<const R *>(reinterpret_cast<const char *>(base) + boff);
}
template <typename R>
__device__ __forceinline__ void sto(R *base, unsigned boff, R v) {
    *reinterpret_cast<R *>(reinterpret_cast<char *>(base) + boff) = v;
}
template <int H, typename R>
__device__ __forceinline__ R DpO(const R *v, unsigned o, long s, const R *hc) {
    R acc = hc[1] * (ldo(v + s, o) - ldo(v, o));
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (ldo(v + k * s, o) - ldo(v - (k - 1) * s, o));
    return acc;
}
template <int H, typename R>
__device__ __forceinline__ R DmO(const R *v, unsigned o, long s, const R *hc) {
    R acc = hc[1] * (ldo(v, o) - ldo(v - s, o));
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (ldo(v + (k - 1) * s, o) - ldo(v - k * s, o));
    return acc;
}

// Non-temporal variants, for streams a kernel touches once (the cell's own
// values and coefficients): they bypass the cache's retention so that the
// stencil neighbours stay resident.
template <typename R>
__device__ __forceinline__ R ldn(const R *base, unsigned boff) {
    return __builtin_nontemporal_load(
        reinterpret_cast<const R *>(reinterpret_cast<const char *>(base) + boff));
}
template <typename R>
__device__ __forceinline__ void stn(R *base, unsigned boff, R v) {
    __builtin_nontemporal_store(v, reinterpret_cast<R *>(reinterpret_cast<char *>(base) + boff));
}

// Wavefield storage narrower than the compute type (FP16 fields with FP32
// arithmetic): loads widen exactly, stores round to nearest even.
template <typename R, typename S>
struct Store {
    static __device__ __forceinline__ R ld(const S *base, unsigned boff) {
        return (R)ldo(base, boff);
    }
    static __device__ __forceinline__ void st(S *base, unsigned boff, R v) { sto(base, boff, (S)v); }
    static __device__ __forceinline__ R ldn_(const S *base, unsigned boff) {
        return (R)ldn(base, boff);
    }
    static __device__ __forceinline__ void stn_(S *base, unsigned boff, R v) {
        stn(base, boff, (S)v);
    }
};
template <>
struct Store<float, __half> {
    static __device__ __forceinline__ float ld(const __half *base, unsigned boff) {
        return __half2float(ldo(base, boff));
    }
    static __device__ __forceinline__ void st(__half *base, unsigned boff, float v) {
        sto(base, boff, __float2half_rn(v));
    }
    static __device__ __forceinline__ float ldn_(const __half *base, unsigned boff) {
        return __half2float(__ushort_as_half(ldn(reinterpret_cast<const unsigned short *>(base), boff)));
    }
    static __device__ __forceinline__ void stn_(__half *base, unsigned boff, float v) {
        stn(reinterpret_cast<unsigned short *>(base), boff, __half_as_ushort(__float2half_rn(v)));
    }
};
template <int H, typename R, typename S>
__device__ __forceinline__ R DpS(const S *v, unsigned o, long s, const R *hc) {
    using T = Store<R, S>;
    R acc = hc[1] * (T::ld(v + s, o) - T::ld(v, o));
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (T::ld(v + k * s, o) - T::ld(v - (k - 1) * s, o));
    return acc;
}
template <int H, typename R, typename S>
__device__ __forceinline__ R DmS(const S *v, unsigned o, long s, const R *hc) {
    using T = Store<R, S>;
    R acc = hc[1] * (T::ld(v, o) - T::ld(v - s, o));
#pragma unroll
    for (int k = 2; k <= H; ++k) acc = acc + hc[k] * (T::ld(v + (k - 1) * s, o) - T::ld(v - k * s, o));
    return acc;
}

}  // namespace
}  // namespace seis

#define SEIS_SWITCH_H(H, CALL)                   \
    switch (H) {                                 \
        case 1: { constexpr int HH = 1; CALL; } break; \
        case 2: { constexpr int HH = 2; CALL; } break; \
        case 3: { constexpr int HH = 3; CALL; } break; \
        case 4: { constexpr int HH = 4; CALL; } break; \
        case 5: { constexpr int HH = 5; CALL; } break; \
        case 6: { constexpr int HH = 6; CALL; } break; \
        default: break;                          \
    }

