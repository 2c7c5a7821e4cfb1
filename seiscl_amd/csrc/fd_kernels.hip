// Staggered-grid velocity / stress kernels for gfx950.
//
// One work-item owns one interior cell; a 256-lane block covers 64 cells of
// a z row (one wave, a contiguous 256-byte line pair of every field) times
// 4 rows in y (3D) or x (2D).  Blocks are renumbered so that consecutive
// hardware blocks, which the dispatcher deals round-robin to the 8 XCDs,
// get tiles that are far apart, and the tiles one XCD runs back to back are
// neighbours: the halo rows a block reads were fetched into that XCD's L2 by
// its neighbour.
//
// Arithmetic follows the reference kernels term by term (src/update_v3D.cl,
// src/update_s3D.cl, src/update_v2D.cl, src/update_s2D.cl, the acoustic
// src/update_{v,s}2D_acc.cl, src/surface{2D,3D}.cl and the generated
// sources/varsout of src/automatic_kernels.c) with the same operation order;
// the build disables FMA contraction, so results are bit-identical to a
// C restatement compiled the same way.
#include <hip/hip_runtime.h>

#include "fd_common.hpp"
#include "seis_internal.hpp"

namespace seis {

namespace {

// Element i of a field typed R* whose storage is S (FP16 storage: the
// pointer holds __half values).
template <typename R, typename S>
__device__ __forceinline__ R getS(const R *p, long i) {
    return Store<R, S>::ld((const S *)p + i, 0u);
}
template <typename R, typename S>
__device__ __forceinline__ void putS(R *p, long i, R v) {
    Store<R, S>::st((S *)p + i, 0u, v);
}

// ---------------------------------------------------------------------------
// 2D P-SV: src/update_v2D.cl:32-246, src/update_s2D.cl:30-326
// ---------------------------------------------------------------------------
//
// S is the storage type of the wavefields and memory variables (R, or
// __half with R = float); coefficients and CPML memories are always R.
template <typename R, typename S, int H>
__device__ __forceinline__ void cell_v2d(const FDArgs<R> &a, int z, int x) {
    using T = Store<R, S>;
    const Layout &g = a.g;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const unsigned os = (unsigned)(i * sizeof(S));
    const R *hc = a.hc;
    R sxx_x = DpS<H>((const S *)a.sxx, os, sx, hc);
    R szz_z = DpS<H>((const S *)a.szz, os, 1, hc);
    R sxz_z = DmS<H>((const S *)a.sxz, os, 1, hc);
    R sxz_x = DmS<H>((const S *)a.sxz, os, sx, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const int k = z > NZ - NAB - 1 ? z - NZ + 2 * NAB : z;
            const bool bot = z > NZ - NAB - 1;
            const int ki = bot ? 2 * NAB - k : k, kh = bot ? 2 * NAB - 1 - k : k;
            const long ip = (long)x * 2 * NAB + k;
            pml(sxz_z, a.psi_sxz_z + ip, cp(a, 2, CB, ki), cp(a, 2, CA, ki), cp(a, 2, CK, ki));
            pml(szz_z, a.psi_szz_z + ip, cp(a, 2, CBH, kh), cp(a, 2, CAH, kh), cp(a, 2, CKH, kh));
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            const long ip = (long)ii * NZ + z;
            pml(sxx_x, a.psi_sxx_x + ip, cp(a, 0, CBH, iH), cp(a, 0, CAH, iH), cp(a, 0, CKH, iH));
            pml(sxz_x, a.psi_sxz_x + ip, cp(a, 0, CB, iI), cp(a, 0, CA, iI), cp(a, 0, CK, iI));
        }
    }
    R vx = T::ld((const S *)a.vx, os) + ((sxx_x + sxz_z) * a.rip[i]);
    R vz = T::ld((const S *)a.vz, os) + ((szz_z + sxz_x) * a.rkp[i]);
    if (a.abs_type == 2) {
        const R *t = a.taper;
        if (!a.freesurf && z < NAB) { vx *= t[z]; vz *= t[z]; }
        if (z > NZ - NAB - 1) { vx *= t[NZ - z - 1]; vz *= t[NZ - z - 1]; }
        if (x < NAB) { vx *= t[x]; vz *= t[x]; }
        if (x > NX - NAB - 1) { vx *= t[NX - x - 1]; vz *= t[NX - x - 1]; }
    }
    T::st((S *)a.vx, os, vx);
    T::st((S *)a.vz, os, vz);
}

template <typename R, typename S, int H>
__global__ __launch_bounds__(256) void k_update_v2d(FDArgs<R> a, int nbz, unsigned nb) {
    unsigned tile;
    xcd_tile(nb, tile);
    if (tile >= nb) return;
    const int bz = tile % nbz, bx = tile / nbz;
    const int z = bz * BZ + threadIdx.x, x = bx * BT + threadIdx.y;
    if (z >= a.g.NZ || x >= a.g.NX) return;
    cell_v2d<R, S, H>(a, z, x);
}

template <typename R, typename S, int H, int L>
__device__ __forceinline__ void cell_s2d(const FDArgs<R> &a, int z, int x) {
    using T = Store<R, S>;
    const Layout &g = a.g;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const unsigned os = (unsigned)(i * sizeof(S));
    const R *hc = a.hc;
    R vxx = DmS<H>((const S *)a.vx, os, sx, hc);
    R vxz = DpS<H>((const S *)a.vx, os, 1, hc);
    R vzz = DmS<H>((const S *)a.vz, os, 1, hc);
    R vzx = DpS<H>((const S *)a.vz, os, sx, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int ki = bot ? 2 * NAB - k : k, kh = bot ? 2 * NAB - 1 - k : k;
            const long ip = (long)x * 2 * NAB + k;
            pml(vxz, a.psi_vx_z + ip, cp(a, 2, CBH, kh), cp(a, 2, CAH, kh), cp(a, 2, CKH, kh));
            pml(vzz, a.psi_vz_z + ip, cp(a, 2, CB, ki), cp(a, 2, CA, ki), cp(a, 2, CK, ki));
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            const long ip = (long)ii * NZ + z;
            pml(vxx, a.psi_vx_x + ip, cp(a, 0, CB, iI), cp(a, 0, CA, iI), cp(a, 0, CK, iI));
            pml(vzx, a.psi_vz_x + ip, cp(a, 0, CBH, iH), cp(a, 0, CAH, iH), cp(a, 0, CKH, iH));
        }
    }
    S *sxzp = (S *)a.sxz, *sxxp = (S *)a.sxx, *szzp = (S *)a.szz;
    R sxz = T::ld(sxzp, os), sxx = T::ld(sxxp, os), szz = T::ld(szzp, os);
    if constexpr (L == 0) {
        const R fipkp = a.muipkp[i];
        const R f = (R)2 * a.mu[i];
        const R gg = a.M[i];
        sxz += (fipkp * (vxz + vzx));
        sxx += (gg * (vxx + vzz)) - (f * vzz);
        szz += (gg * (vxx + vzz)) - (f * vxx);
    } else {
        const R DT2 = a.DT2;
        const R gg = a.cg[i], f = a.cf[i], d = a.cd[i], e = a.ce[i];
        const R fipkp = a.fipkp[i], dipkp = a.dipkp[i];
        const long nf = g.nfield;
        R rxz[L], rxx[L], rzz[L];
        R sumrxz = 0, sumrxx = 0, sumrzz = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const long ln = l * nf;
            rxz[l] = T::ld((const S *)a.rxz + ln, os);
            rxx[l] = T::ld((const S *)a.rxx + ln, os);
            rzz[l] = T::ld((const S *)a.rzz + ln, os);
            sumrxz += rxz[l]; sumrxx += rxx[l]; sumrzz += rzz[l];
        }
        const R lsxz = (fipkp * (vxz + vzx)) + (DT2 * sumrxz);
        const R lsxx = ((gg * (vxx + vzz)) - (f * vzz)) + (DT2 * sumrxx);
        const R lszz = ((gg * (vxx + vzz)) - (f * vxx)) + (DT2 * sumrzz);
        sumrxz = sumrxx = sumrzz = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const R b = a.bl[l], c = a.cl[l], et = a.eta[l];
            const long ln = l * nf;
            rxz[l] = b * (rxz[l] * c - et * (dipkp * (vxz + vzx)));
            rxx[l] = b * (rxx[l] * c - et * ((e * (vxx + vzz)) - (d * vzz)));
            rzz[l] = b * (rzz[l] * c - et * ((e * (vxx + vzz)) - (d * vxx)));
            T::st((S *)a.rxz + ln, os, rxz[l]);
            T::st((S *)a.rxx + ln, os, rxx[l]);
            T::st((S *)a.rzz + ln, os, rzz[l]);
            sumrxz += rxz[l]; sumrxx += rxx[l]; sumrzz += rzz[l];
        }
        sxz += lsxz + (DT2 * sumrxz);
        sxx += lsxx + (DT2 * sumrxx);
        szz += lszz + (DT2 * sumrzz);
    }
    if (a.abs_type == 2) {
        const R *t = a.taper;
        if (!a.freesurf && z < NAB) { sxx *= t[z]; szz *= t[z]; sxz *= t[z]; }
        if (z > NZ - NAB - 1) { sxx *= t[NZ - z - 1]; szz *= t[NZ - z - 1]; sxz *= t[NZ - z - 1]; }
        if (x < NAB) { sxx *= t[x]; szz *= t[x]; sxz *= t[x]; }
        if (x > NX - NAB - 1) { sxx *= t[NX - x - 1]; szz *= t[NX - x - 1]; sxz *= t[NX - x - 1]; }
    }
    T::st(sxzp, os, sxz);
    T::st(sxxp, os, sxx);
    T::st(szzp, os, szz);
}

template <typename R, typename S, int H, int L>
__global__ __launch_bounds__(256) void k_update_s2d(FDArgs<R> a, int nbz, unsigned nb) {
    unsigned tile;
    xcd_tile(nb, tile);
    if (tile >= nb) return;
    const int bz = tile % nbz, bx = tile / nbz;
    const int z = bz * BZ + threadIdx.x, x = bx * BT + threadIdx.y;
    if (z >= a.g.NZ || x >= a.g.NX) return;
    cell_s2d<R, S, H, L>(a, z, x);
}

// ---------------------------------------------------------------------------
// 2D acoustic: src/update_v2D_acc.cl:80-325, src/update_s2D_acc.cl:79-402
// ---------------------------------------------------------------------------
template <typename R, typename S, int H>
__device__ __forceinline__ void cell_v2d_acc(const FDArgs<R> &a, int z, int x) {
    using T = Store<R, S>;
    const Layout &g = a.g;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const unsigned os = (unsigned)(i * sizeof(S));
    const R *hc = a.hc;
    R p_x = DpS<H>((const S *)a.p, os, sx, hc);
    R p_z = DpS<H>((const S *)a.p, os, 1, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int kh = bot ? 2 * NAB - 1 - k : k;
            pml(p_z, a.psi_p_z + (long)x * 2 * NAB + k, cp(a, 2, CBH, kh), cp(a, 2, CAH, kh),
                cp(a, 2, CKH, kh));
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            pml(p_x, a.psi_p_x + (long)ii * NZ + z, cp(a, 0, CBH, iH), cp(a, 0, CAH, iH),
                cp(a, 0, CKH, iH));
        }
    }
    R vx = T::ld((const S *)a.vx, os) + (p_x * a.rip[i]);
    R vz = T::ld((const S *)a.vz, os) + (p_z * a.rkp[i]);
    if (a.abs_type == 2) {
        const R *t = a.taper;
        // the acoustic kernel tapers the top edge even with a free surface
        if (z < NAB) { vx *= t[z]; vz *= t[z]; }
        if (z > NZ - NAB - 1) { vx *= t[NZ - z - 1]; vz *= t[NZ - z - 1]; }
        if (x < NAB) { vx *= t[x]; vz *= t[x]; }
        if (x > NX - NAB - 1) { vx *= t[NX - x - 1]; vz *= t[NX - x - 1]; }
    }
    T::st((S *)a.vx, os, vx);
    T::st((S *)a.vz, os, vz);
}

template <typename R, typename S, int H>
__global__ __launch_bounds__(256) void k_update_v2d_acc(FDArgs<R> a, int nbz, unsigned nb) {
    unsigned tile;
    xcd_tile(nb, tile);
    if (tile >= nb) return;
    const int bz = tile % nbz, bx = tile / nbz;
    const int z = bz * BZ + threadIdx.x, x = bx * BT + threadIdx.y;
    if (z >= a.g.NZ || x >= a.g.NX) return;
    cell_v2d_acc<R, S, H>(a, z, x);
}

template <typename R, typename S, int H, int L>
__device__ __forceinline__ void cell_s2d_acc(const FDArgs<R> &a, int z, int x) {
    using T = Store<R, S>;
    const Layout &g = a.g;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const unsigned os = (unsigned)(i * sizeof(S));
    const R *hc = a.hc;
    R vxx = DmS<H>((const S *)a.vx, os, sx, hc);
    R vzz = DmS<H>((const S *)a.vz, os, 1, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int ki = bot ? 2 * NAB - k : k;
            pml(vzz, a.psi_vz_z + (long)x * 2 * NAB + k, cp(a, 2, CB, ki), cp(a, 2, CA, ki),
                cp(a, 2, CK, ki));
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii;
            pml(vxx, a.psi_vx_x + (long)ii * NZ + z, cp(a, 0, CB, iI), cp(a, 0, CA, iI),
                cp(a, 0, CK, iI));
        }
    }
    R p = T::ld((const S *)a.p, os);
    if constexpr (L == 0) {
        p += (a.M[i] * (vxx + vzz));
    } else {
        const R DT2 = a.DT2, gg = a.cg[i], e = a.ce[i];
        const long nf = g.nfield;
        R rp[L];
        R sumrp = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) { rp[l] = T::ld((const S *)a.rp + l * nf, os); sumrp += rp[l]; }
        const R lp = (gg * (vxx + vzz)) + (DT2 * sumrp);
        sumrp = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            rp[l] = a.bl[l] * (rp[l] * a.cl[l] - a.eta[l] * e * (vxx + vzz));
            T::st((S *)a.rp + l * nf, os, rp[l]);
            sumrp += rp[l];
        }
        p += lp + (DT2 * sumrp);
    }
    if (a.abs_type == 2) {
        const R *t = a.taper;
        if (z < NAB) p *= t[z];
        if (z > NZ - NAB - 1) p *= t[NZ - z - 1];
        if (x < NAB) p *= t[x];
        if (x > NX - NAB - 1) p *= t[NX - x - 1];
    }
    T::st((S *)a.p, os, p);
}

template <typename R, typename S, int H, int L>
__global__ __launch_bounds__(256) void k_update_s2d_acc(FDArgs<R> a, int nbz, unsigned nb) {
    unsigned tile;
    xcd_tile(nb, tile);
    if (tile >= nb) return;
    const int bz = tile % nbz, bx = tile / nbz;
    const int z = bz * BZ + threadIdx.x, x = bx * BT + threadIdx.y;
    if (z >= a.g.NZ || x >= a.g.NX) return;
    cell_s2d_acc<R, S, H, L>(a, z, x);
}

// ---------------------------------------------------------------------------
// Free surface by the image method (Levander 1988): src/surface2D.cl:30-193,
// src/surface3D.cl:39-271.  One lane per surface column.
// ---------------------------------------------------------------------------
template <typename R, typename S, int H>
__global__ void k_surface2d(FDArgs<R> a, int pdir) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const Layout &g = a.g;
    if (x >= g.NX) return;
    const long i = g.idx(0, 0, x), sx = g.sx, nf = g.nfield;
    const int L = a.L, NAB = a.nab, NX = g.NX, NZ = g.NZ;
    R *szz = a.szz, *sxz = a.sxz, *sxx = a.sxx, *rzz = a.rzz, *rxx = a.rxx;
    putS<R, S>(szz, i, 0);
    for (int l = 0; l < L; ++l) putS<R, S>(rzz, l * nf + i, 0);
    for (int m = 1; m <= H; ++m) {
        putS<R, S>(szz, i - m, -getS<R, S>(szz, i + m));
        putS<R, S>(sxz, i - m, -getS<R, S>(sxz, i + m - 1));
    }
    const unsigned os = (unsigned)(i * sizeof(S));
    R vxx = DmS<H>((const S *)a.vx, os, sx, a.hc);
    R vzz = DmS<H>((const S *)a.vz, os, 1, a.hc);
    if (a.abs_type == 1) {
        if (x < NAB) vxx = vxx / cp(a, 0, CK, x) + a.psi_vx_x[(long)x * NZ];
        if (x > NX - NAB - 1) {
            const int ii = x - NX + 2 * NAB;
            vxx = vxx / cp(a, 0, CK, 2 * NAB - ii) + a.psi_vx_x[(long)ii * NZ];
        }
    }
    const R fp = (R)pdir;
    if (L == 0) {
        const R f = a.mu[i] * (R)2, gg = a.M[i];
        const R h = -((gg - f) * (gg - f) * (vxx) / gg) - ((gg - f) * vzz);
        putS<R, S>(sxx, i, getS<R, S>(sxx, i) + fp * h);
    } else {
        const R f = a.cf[i], gg = a.cg[i], d = a.cd[i], e = a.ce[i], DT2 = a.DT2;
        R h = -((gg - f) * (gg - f) * (vxx) / gg) - ((gg - f) * vzz);
        R sump = 0;
        for (int l = 0; l < L; ++l) sump += getS<R, S>(rxx, l * nf + i);
        h += -DT2 * sump;
        R sxxv = getS<R, S>(sxx, i) + fp * h;
        sump = 0;
        for (int l = 0; l < L; ++l) {
            // b = eta/(1+eta/2) and the (f/g - 1.0) factor are doubles in
            // the reference expression
            const double b = (double)a.eta[l] / (1.0 + ((double)a.eta[l] * 0.5));
            const R bb = (R)b;
            const R hh = (R)((double)bb * ((double)(d - e) * ((double)(f / gg) - 1.0) * (double)vxx -
                                          (double)((d - e) * vzz)));
            const R r = getS<R, S>(rxx, l * nf + i) + fp * hh;
            putS<R, S>(rxx, l * nf + i, r);
            sump += getS<R, S>(rxx, l * nf + i);
        }
        putS<R, S>(sxx, i, sxxv + fp * (DT2 * sump));
    }
}

template <typename R, int H>
__global__ void k_surface3d(FDArgs<R> a, int pdir) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    const Layout &g = a.g;
    if (y >= g.NY) return;
    const long i = g.idx(0, y, x), sx = g.sx, sy = g.sy, nf = g.nfield;
    const int L = a.L, NAB = a.nab, NX = g.NX, NY = g.NY, NZ = g.NZ;
    a.szz[i] = 0;
    for (int l = 0; l < L; ++l) a.rzz[l * nf + i] = 0;
    for (int m = 1; m <= H; ++m) {
        a.szz[i - m] = -a.szz[i + m];
        a.sxz[i - m] = -a.sxz[i + m - 1];
        a.syz[i - m] = -a.syz[i + m - 1];
    }
    R vxx = Dm<H>(a.vx, i, sx, a.hc);
    R vyy = Dm<H>(a.vy, i, sy, a.hc);
    R vzz = Dm<H>(a.vz, i, 1, a.hc);
    if (a.abs_type == 1) {
        if (y < NAB) vyy = vyy / cp(a, 1, CK, y) + a.psi_vy_y[(long)x * 2 * NAB * NZ + (long)y * NZ];
        else if (y > NY - NAB - 1) {
            const int j = y - NY + 2 * NAB;
            vyy = vyy / cp(a, 1, CK, 2 * NAB - j) + a.psi_vy_y[(long)x * 2 * NAB * NZ + (long)j * NZ];
        }
        if (x < NAB) vxx = vxx / cp(a, 0, CK, x) + a.psi_vx_x[(long)x * NY * NZ + (long)y * NZ];
        if (x > NX - NAB - 1) {
            const int ii = x - NX + 2 * NAB;
            vxx = vxx / cp(a, 0, CK, 2 * NAB - ii) + a.psi_vx_x[(long)ii * NY * NZ + (long)y * NZ];
        }
    }
    if (L == 0) {
        const R f = a.mu[i] * (R)2, gg = a.M[i];
        const R h = -((gg - f) * (gg - f) * (vxx + vyy) / gg) - ((gg - f) * vzz);
        a.sxx[i] += h;   // the 3D elastic kernel does not apply pdir
        a.syy[i] += h;
    } else {
        const R fp = (R)pdir;
        const R f = a.cf[i], gg = a.cg[i], d = a.cd[i], e = a.ce[i], DT2 = a.DT2;
        const R h = -((gg - f) * (gg - f) * (vxx + vyy) / gg) - ((gg - f) * vzz);
        R sumx = 0, sumy = 0;
        for (int l = 0; l < L; ++l) { sumx += a.rxx[l * nf + i]; sumy += a.ryy[l * nf + i]; }
        a.sxx[i] += fp * (h - DT2 * sumx);
        a.syy[i] += fp * (h - DT2 * sumy);
        sumx = 0;
        sumy = 0;
        for (int l = 0; l < L; ++l) {
            const double b = (double)a.eta[l] / (1.0 + ((double)a.eta[l] * 0.5));
            const R bb = (R)b;
            const R hh = (R)((double)bb * ((double)(d - e) * ((double)(f / gg) - 1.0) * (double)(vxx + vyy) -
                                          (double)((d - e) * vzz)));
            a.rxx[l * nf + i] += hh;
            a.ryy[l * nf + i] += hh;
            sumx += a.rxx[l * nf + i];
            sumy += a.ryy[l * nf + i];
        }
        a.sxx[i] += (DT2 * sumx);
        a.syy[i] += (DT2 * sumy);
    }
}

// ---------------------------------------------------------------------------
// Sources (src/automatic_kernels.c:394-629) and receivers (:22-219).
// One lane walks all sources in order, so sources sharing a cell add up in
// a fixed order.
// ---------------------------------------------------------------------------

// src holds each source already multiplied by the storage scale of the
// field it feeds (1 unless wavefields are stored scaled).
template <typename R, typename S>
__device__ void sources_body(const FDArgs<R> &a, int nd_case, const ShotDev &s, int NT, int nt,
                             int pdir) {
    for (int k = 0; k < s.nsrc; ++k) {
        const int cell = s.src_cell[k];
        if (cell < 0) continue;
        if (a.freesurf && pdir == -1 && s.src_k[k] < 2 * a.g.H) continue;
        const int type = s.src_type[k];
        const R amp = (R)pdir * ((R)a.DT * (R)s.src[(long)k * NT + nt]);
        R *f = nullptr;
        if (type == 0) f = a.vx;
        else if (type == 1) f = a.vy;
        else if (type == 2) f = a.vz;
        else if (type != 3 && type != 100) continue;
        else if (nd_case == 22) f = a.p;
        if (f) {
            putS<R, S>(f, cell, getS<R, S>(f, cell) + amp);
            continue;
        }
        const double nd = nd_case == 3 ? 3.0 : 2.0;   // "+=amp/%f" (double)
        R *st[3] = {a.sxx, a.szz, nd_case == 3 ? a.syy : nullptr};
        for (int q = 0; q < 3; ++q)
            if (st[q]) putS<R, S>(st[q], cell, (R)((double)getS<R, S>(st[q], cell) + (double)amp / nd));
    }
}

// tbase (may be NULL): a device step counter added to nt, so that a captured
// sequence of steps can be replayed at any starting step
template <typename R, typename S>
__global__ void k_sources(FDArgs<R> a, int nd_case, ShotDev s, int NT, int nt, int pdir,
                          const int *tbase) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sources_body<R, S>(a, nd_case, s, NT, nt + (tbase ? *tbase : 0), pdir);
}

struct OutList {
    int n;
    int code[10];
    float unscale[10];   // inverse storage scale of each output
};

template <typename R, typename S>
__device__ void varsout_body(const FDArgs<R> &a, int nd_case, const ShotDev &s, const OutList &o,
                             int nt, int r) {
    const int cell = s.rec_cell[r];
    if (cell < 0) return;
    float *out = s.out + (long)nt * o.n * s.nrec;
    for (int q = 0; q < o.n; ++q) {
        R v;
        switch (o.code[q]) {
            case O_VX: v = getS<R, S>(a.vx, cell); break;
            case O_VY: v = getS<R, S>(a.vy, cell); break;
            case O_VZ: v = getS<R, S>(a.vz, cell); break;
            case O_SXX: v = getS<R, S>(a.sxx, cell); break;
            case O_SYY: v = getS<R, S>(a.syy, cell); break;
            case O_SZZ: v = getS<R, S>(a.szz, cell); break;
            case O_SXZ: v = getS<R, S>(a.sxz, cell); break;
            case O_SXY: v = getS<R, S>(a.sxy, cell); break;
            case O_SYZ: v = getS<R, S>(a.syz, cell); break;
            default:
                if (nd_case == 22) v = getS<R, S>(a.p, cell);
                else {
                    R sm = getS<R, S>(a.sxx, cell) + getS<R, S>(a.szz, cell);
                    if (nd_case == 3) sm = sm + getS<R, S>(a.syy, cell);
                    v = (R)((double)sm / (nd_case == 3 ? 3.0 : 2.0));
                }
        }
        if constexpr (sizeof(R) == 8)
            ((double *)s.out)[((long)nt * o.n + q) * s.nrec + r] = v * (R)o.unscale[q];
        else
            out[(long)q * s.nrec + r] = (float)v * o.unscale[q];
    }
}

template <typename R, typename S>
__global__ void k_varsout(FDArgs<R> a, int nd_case, ShotDev s, OutList o, int nt,
                          const int *tbase) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.nrec) return;
    varsout_body<R, S>(a, nd_case, s, o, nt + (tbase ? *tbase : 0), r);
}

// Receivers of step t_rec, then the sources of step t_src, in one launch:
// nothing runs between varsout of step t-1 and the sources of step t
// (src/time_stepping.c:785-816), so the two share a launch.  One block, so
// that every receiver has been read before a source writes (a receiver may
// sit on a source cell).  Steps outside [0, NT) are skipped; tbase as above.
template <typename R, typename S>
__global__ __launch_bounds__(1024) void k_io(FDArgs<R> a, int nd_case, ShotDev s, OutList o,
                                              int NT, int t_rec, int t_src, const int *tbase) {
    const int tb = tbase ? *tbase : 0;
    const int tr = t_rec + tb, ts = t_src + tb;
    if (tr >= 0 && tr < NT)
        for (int r = threadIdx.x; r < s.nrec; r += blockDim.x) varsout_body<R, S>(a, nd_case, s, o, tr, r);
    __syncthreads();
    if (threadIdx.x == 0 && ts >= 0 && ts < NT) sources_body<R, S>(a, nd_case, s, NT, ts, 1);
}

// the step counter of replayed step sequences: set (add = 0) or advance
__global__ void k_tcounter(int *t, int v, int add) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t = add ? *t + v : v;
}

// ---------------------------------------------------------------------------
// Batched time steps (serving): every launch advances all active slots of a
// batch, each slot a whole modeling job at its own time step.  The per-cell
// arithmetic is the single-shot kernels' (the same device functions), so a
// slot's results do not depend on what else is in the batch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int batch_find(const unsigned *tile0, int n, unsigned tile) {
    int lo = 0, hi = n;           // largest k with tile0[k] <= tile
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tile0[mid] <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

// KIND 0: elastic velocity, 1: elastic stress, 2: acoustic velocity, 3: acoustic pressure
template <int H, int KIND>
__device__ __forceinline__ void batch_cell(const FDArgs<float> &a, int z, int x) {
    if constexpr (KIND == 0) cell_v2d<float, float, H>(a, z, x);
    else if constexpr (KIND == 1) cell_s2d<float, float, H, 0>(a, z, x);
    else if constexpr (KIND == 2) cell_v2d_acc<float, float, H>(a, z, x);
    else cell_s2d_acc<float, float, H, 0>(a, z, x);
}

// The tiles [lt0, lt0 + kBatchTiles) of one slot, in its z-fastest tile order.
template <int H, int KIND>
__device__ __forceinline__ void batch_tiles(const FDArgs<float> &a, unsigned lt0, int tiles) {
    const int NZ = a.g.NZ, NX = a.g.NX;
    const unsigned nbz = (NZ + BZ - 1) / BZ, nt = nbz * ((NX + BT - 1) / BT);
#pragma unroll 1
    for (int i = 0; i < tiles; ++i) {
        const unsigned lt = lt0 + i;
        if (lt >= nt) break;
        const int z = (lt % nbz) * BZ + threadIdx.x, x = (lt / nbz) * BT + threadIdx.y;
        if (z < NZ && x < NX) batch_cell<H, KIND>(a, z, x);
    }
}

// One pass of one batch step over every slot: PASS 0 the velocities, 1 the
// stresses (pressure).  A block finds its slot by binary search in the block
// prefix and updates kBatchTiles tiles of it, which spreads the cost of the
// search and of reading the slot over more cells.  The slot's kind and FD
// order select the cell update (uniform per block); slots are grouped by kind
// and order, so neighbouring blocks run the same code.
template <int PASS>
__global__ __launch_bounds__(256) void k_batch_pass(const BatchSlot *__restrict__ slots,
                                                    BatchTable tab, int step) {
    unsigned blk;
    xcd_tile(tab.nblocks, blk);
    if (blk >= tab.nblocks) return;
    const int k = batch_find(tab.block0, tab.n, blk);
    const BatchSlot &S = slots[tab.ids[k]];
    const int t = step - S.start;
    if (!S.active || t < 0 || t >= S.tmax) return;
    const unsigned lt0 = (blk - tab.block0[k]) * tab.tiles;
    if (S.nd_case == 22) {
        SEIS_SWITCH_H(S.a.g.H, (batch_tiles<HH, PASS + 2>(S.a, lt0, tab.tiles)));
    } else {
        SEIS_SWITCH_H(S.a.g.H, (batch_tiles<HH, PASS>(S.a, lt0, tab.tiles)));
    }
}

__global__ void k_batch_sources(const BatchSlot *slots, int n, int step) {
    if (threadIdx.x != 0 || (int)blockIdx.x >= n) return;
    const BatchSlot &S = slots[blockIdx.x];
    const int t = step - S.start;
    if (!S.active || t < 0 || t >= S.tmax) return;
    sources_body<float, float>(S.a, S.nd_case, S.shot, S.NT, t, 1);
}

__global__ void k_batch_varsout(const BatchSlot *slots, int n, int step) {
    if ((int)blockIdx.y >= n) return;
    const BatchSlot &S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = step - S.start;
    if (!S.active || t < 0 || t >= S.tmax || r >= S.shot.nrec) return;
    OutList o;
    o.n = S.nout;
    for (int q = 0; q < S.nout; ++q) {
        o.code[q] = S.codes[q];
        o.unscale[q] = 1.0f;
    }
    varsout_body<float, float>(S.a, S.nd_case, S.shot, o, t, r);
}

__global__ void k_fill(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}

template <typename R, typename S, int H>
void dispatch_v(const FDArgs<R> &a, int nd_case, hipStream_t st) {
    const Layout &g = a.g;
    const int nbz = (g.NZ + BZ - 1) / BZ;
    dim3 blk(BZ, BT);
    {
        const int nbx = (g.NX + BT - 1) / BT;
        const unsigned nb = (unsigned)nbz * nbx;
        if (nd_case == 22)
            hipLaunchKernelGGL((k_update_v2d_acc<R, S, H>), dim3(((nb + 7) / 8) * 8), blk, 0, st, a, nbz, nb);
        else
            hipLaunchKernelGGL((k_update_v2d<R, S, H>), dim3(((nb + 7) / 8) * 8), blk, 0, st, a, nbz, nb);
    }
}

template <typename R, typename S, int H, int L>
void dispatch_s_L(const FDArgs<R> &a, int nd_case, hipStream_t st) {
    const Layout &g = a.g;
    const int nbz = (g.NZ + BZ - 1) / BZ;
    dim3 blk(BZ, BT);
    {
        const int nbx = (g.NX + BT - 1) / BT;
        const unsigned nb = (unsigned)nbz * nbx;
        if (nd_case == 22)
            hipLaunchKernelGGL((k_update_s2d_acc<R, S, H, L>), dim3(((nb + 7) / 8) * 8), blk, 0, st, a, nbz, nb);
        else
            hipLaunchKernelGGL((k_update_s2d<R, S, H, L>), dim3(((nb + 7) / 8) * 8), blk, 0, st, a, nbz, nb);
    }
}

template <typename R, typename S, int H>
void dispatch_s(const FDArgs<R> &a, int nd_case, hipStream_t st) {
    switch (a.L) {
        case 0: dispatch_s_L<R, S, H, 0>(a, nd_case, st); break;
        case 1: dispatch_s_L<R, S, H, 1>(a, nd_case, st); break;
        case 2: dispatch_s_L<R, S, H, 2>(a, nd_case, st); break;
        case 3: dispatch_s_L<R, S, H, 3>(a, nd_case, st); break;
        case 4: dispatch_s_L<R, S, H, 4>(a, nd_case, st); break;
        default: break;   // rejected at model creation
    }
}

template <typename R, typename S, int H>
void dispatch_surface(const FDArgs<R> &a, int nd_case, int pdir, hipStream_t st) {
    const Layout &g = a.g;
    if (nd_case == 3) {
        // 3D runs with FP16 storage fuse the surface into the stress update
        if constexpr (std::is_same<R, S>::value)
            hipLaunchKernelGGL((k_surface3d<R, H>), dim3((g.NY + 63) / 64, g.NX), dim3(64), 0, st,
                               a, pdir);
    } else {
        hipLaunchKernelGGL((k_surface2d<R, S, H>), dim3((g.NX + 63) / 64), dim3(64), 0, st, a,
                           pdir);
    }
}

}  // namespace

template <typename R, typename S>
void launch_update_v(const FDArgs<R> &a, int nd_case, void *stream) {
    if (nd_case == 3) return launch_v3d<R, S>(a, stream);
    SEIS_SWITCH_H(a.g.H, (dispatch_v<R, S, HH>(a, nd_case, (hipStream_t)stream)));
}
template <typename R, typename S>
void launch_update_s(const FDArgs<R> &a, int nd_case, void *stream) {
    if (nd_case == 3) return launch_s3d<R, S>(a, stream);
    SEIS_SWITCH_H(a.g.H, (dispatch_s<R, S, HH>(a, nd_case, (hipStream_t)stream)));
}
template <typename R, typename S>
void launch_surface(const FDArgs<R> &a, int nd_case, int pdir, void *stream) {
    SEIS_SWITCH_H(a.g.H, (dispatch_surface<R, S, HH>(a, nd_case, pdir, (hipStream_t)stream)));
}
template <typename R, typename S>
void launch_sources(const FDArgs<R> &a, int nd_case, const ShotDev &s, int NT, int nt,
                    int pdir, void *stream, const int *tbase) {
    if (s.nsrc <= 0) return;
    hipLaunchKernelGGL((k_sources<R, S>), dim3(1), dim3(64), 0, (hipStream_t)stream, a, nd_case,
                       s, NT, nt, pdir, tbase);
}
template <typename R, typename S>
void launch_varsout(const FDArgs<R> &a, int nd_case, const ShotDev &s, const int *codes,
                    const float *unscale, int nout, int nt, void *stream, const int *tbase) {
    if (s.nrec <= 0) return;
    OutList o;
    o.n = nout;
    for (int q = 0; q < nout && q < 10; ++q) {
        o.code[q] = codes[q];
        o.unscale[q] = unscale ? unscale[q] : 1.0f;
    }
    hipLaunchKernelGGL((k_varsout<R, S>), dim3((s.nrec + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, a, nd_case, s, o, nt, tbase);
}

template <typename R, typename S>
void launch_io(const FDArgs<R> &a, int nd_case, const ShotDev &s, const int *codes,
               const float *unscale, int nout, int NT, int t_rec, int t_src, void *stream,
               const int *tbase) {
    if (s.nrec <= 0 && s.nsrc <= 0) return;
    OutList o;
    o.n = nout;
    for (int q = 0; q < nout && q < 10; ++q) {
        o.code[q] = codes[q];
        o.unscale[q] = unscale ? unscale[q] : 1.0f;
    }
    hipLaunchKernelGGL((k_io<R, S>), dim3(1), dim3(1024), 0, (hipStream_t)stream, a, nd_case, s,
                       o, NT, t_rec, t_src, tbase);
}

void launch_tcounter(int *t, int v, int add, void *stream) {
    hipLaunchKernelGGL(k_tcounter, dim3(1), dim3(64), 0, (hipStream_t)stream, t, v, add);
}

void launch_batch_steps(const BatchSlot *dslots, int nslots, const BatchTable &tab, int maxrec,
                        int step0, int nsteps, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    if (nslots <= 0) return;
    const dim3 grid(((tab.nblocks + 7) / 8) * 8), blk(BZ, BT);
    for (int k = 0; k < nsteps; ++k) {
        const int step = step0 + k;
        hipLaunchKernelGGL(k_batch_sources, dim3(nslots), dim3(64), 0, st, dslots, nslots, step);
        if (tab.nblocks) {
            hipLaunchKernelGGL(k_batch_pass<0>, grid, blk, 0, st, dslots, tab, step);
            hipLaunchKernelGGL(k_batch_pass<1>, grid, blk, 0, st, dslots, tab, step);
        }
        if (maxrec > 0)
            hipLaunchKernelGGL(k_batch_varsout, dim3((maxrec + 255) / 256, nslots), dim3(256), 0,
                               st, dslots, nslots, step);
    }
}

void launch_fill(void *ptr, size_t bytes, void *stream) {
    if (!ptr || bytes == 0) return;
    const size_t n = bytes / 4;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (uint32_t *)ptr, n);
}

template void launch_update_v<float, float>(const FDArgs<float> &, int, void *);
template void launch_update_s<float, float>(const FDArgs<float> &, int, void *);
template void launch_surface<float, float>(const FDArgs<float> &, int, int, void *);
template void launch_update_v<float, __half>(const FDArgs<float> &, int, void *);
template void launch_update_s<float, __half>(const FDArgs<float> &, int, void *);
template void launch_surface<float, __half>(const FDArgs<float> &, int, int, void *);
template void launch_update_v<double, double>(const FDArgs<double> &, int, void *);
template void launch_update_s<double, double>(const FDArgs<double> &, int, void *);
template void launch_surface<double, double>(const FDArgs<double> &, int, int, void *);
template void launch_sources<float, float>(const FDArgs<float> &, int, const ShotDev &, int, int,
                                          int, void *, const int *);
template void launch_sources<float, __half>(const FDArgs<float> &, int, const ShotDev &, int, int,
                                           int, void *, const int *);
template void launch_sources<double, double>(const FDArgs<double> &, int, const ShotDev &, int,
                                            int, int, void *, const int *);
template void launch_io<float, float>(const FDArgs<float> &, int, const ShotDev &, const int *,
                                     const float *, int, int, int, int, void *, const int *);
template void launch_io<float, __half>(const FDArgs<float> &, int, const ShotDev &, const int *,
                                      const float *, int, int, int, int, void *, const int *);
template void launch_io<double, double>(const FDArgs<double> &, int, const ShotDev &,
                                       const int *, const float *, int, int, int, int, void *,
                                       const int *);
template void launch_varsout<float, float>(const FDArgs<float> &, int, const ShotDev &,
                                          const int *, const float *, int, int, void *,
                                          const int *);
template void launch_varsout<float, __half>(const FDArgs<float> &, int, const ShotDev &,
                                           const int *, const float *, int, int, void *,
                                           const int *);
template void launch_varsout<double, double>(const FDArgs<double> &, int, const ShotDev &,
                                            const int *, const float *, int, int, void *,
                                            const int *);

}  // namespace seis
