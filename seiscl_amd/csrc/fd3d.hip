// 3D velocity and stress updates.
//
// One lane owns one interior cell; a block covers 64 cells along z (one
// wave reads two whole 128-byte lines of every field) times BY rows in y at a
// single x-plane.  Tiles are numbered z fastest, then x, then the y slab,
// and each XCD receives a contiguous range of that numbering.  The blocks an
// XCD holds at any moment therefore cover consecutive x-planes of one y slab:
// the x-neighbour planes a stencil reads are being read by co-resident
// blocks and are served by that XCD's L2, and a y slab's halo rows are
// re-read only once per slab.
//
// The free surface (image method, src/surface3D.cl:39-271) is applied in the
// stress kernel's epilogue by the lanes that own z = 0..H, on the values they
// just computed; it is skipped when the caller needs the stresses between
// the stress update and the surface (boundary saving for gradients).
//
// Arithmetic: src/update_v3D.cl:34-369 and src/update_s3D.cl:35-489, term by
// term and in the same order (FMA contraction disabled in the build).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "fd_common.hpp"
#include "seis_internal.hpp"

namespace seis {
namespace {

// rows of y per block of the velocity and of the stress update.  The FP32
// stress update reads three y-derivatives and runs at 4 waves per SIMD
// either way, so 8 rows (a 512-lane block) cut its y-halo re-reads (256^3,
// with the non-temporal accesses below: viscoelastic 0.712 ms with 4 rows,
// 0.699 with 8, 0.709 with 16; elastic 0.339, 0.335, 0.359); with FP16
// storage or FP64 arithmetic 4 rows measure as fast or faster (in FP64 a
// 1024-lane block would cap the kernel at 128 VGPRs and spill).  The
// velocity update keeps 4 rows, the block size that lets it run 6 waves.
#ifndef SEIS_BY_V
#define SEIS_BY_V 4
#endif
#ifndef SEIS_BY_S
#define SEIS_BY_S 8
#endif
constexpr int BYV = SEIS_BY_V;
template <typename R, typename S>
constexpr int bys() { return sizeof(R) == 4 && sizeof(S) == 4 ? SEIS_BY_S : 4; }

// Non-temporal access (the nt bit on loads and stores) for the streams each
// update touches exactly once: the cell's own fields, its coefficients, its
// memory variables and the stores.  They then do not displace the stencil
// neighbours that co-resident blocks still read from L2.  Measured at 256^3
// (FD4, CPML, free surface) against plain accesses: viscoelastic L=2 -11%
// per step, elastic -4%, FP16 storage -8%, FP64 -8%.
// Bits, stress update: 1 the cell's own stresses, 2 coefficients, 4 memory
// variables (loads), 8 the stores.
#ifndef SEIS_NT_S
#define SEIS_NT_S 15
#endif
constexpr int kNTS = SEIS_NT_S;
// velocity update: 1 the cell's own velocities, 2 coefficients, 8 the stores
#ifndef SEIS_NT_V
#define SEIS_NT_V 11
#endif
constexpr int kNTV = SEIS_NT_V;

// CPML memory variables of both updates: non-temporal (1) or plain (0).
#ifndef SEIS_NT_PSI
#define SEIS_NT_PSI 1
#endif
constexpr bool kNTP = SEIS_NT_PSI;

// Interleave the y slabs over the XCDs (1) or give each XCD a contiguous
// run of them (0).
#ifndef SEIS_SLAB_IL
#define SEIS_SLAB_IL 1
#endif

// tile -> (bz, y slab, x) with z fastest, then x, then the y slab.  Each XCD
// receives a contiguous run of tiles, i.e. whole y slabs.  When the slab
// count is a multiple of 8 the slabs are dealt out round-robin (XCD k gets
// slabs k, k + 8, ...): a contiguous split would give the two outer XCDs
// every slab of the y-CPML strips, and the kernel lasts as long as its
// slowest XCD.
// With yg > 1 (nby a multiple of yg), yg adjacent slabs form a group whose
// tiles are numbered z fastest, then the slab within the group, then x: the
// blocks an XCD holds at once cover a few x-planes of all the group's slabs,
// so the y-halo rows one slab reads are the rows its neighbour slab is
// reading at the same time.  Groups are dealt to the XCDs as slabs are.
template <int BY>
__device__ __forceinline__ bool tile3(int nbz, int NX, int NY, unsigned nb, int yg, int &z,
                                      int &y, int &x) {
    unsigned tile;
    xcd_tile(nb, tile);
    if (tile >= nb) return false;
    const int bz = tile % nbz;
    unsigned r = tile / nbz;
    const int nby = (NY + BY - 1) / BY;
    int by;
    if (yg > 1) {
        const int sub = r % yg;
        r /= yg;
        x = r % NX;
        int grp = r / NX;
        const int ng = nby / yg;
        if (SEIS_SLAB_IL && ng % 8 == 0) {
            const int m = ng / 8;
            grp = (grp % m) * 8 + grp / m;
        }
        by = grp * yg + sub;
    } else {
        x = r % NX;
        by = r / NX;
        if (SEIS_SLAB_IL && nby % 8 == 0) {
            const int m = nby / 8;
            by = (by % m) * 8 + by / m;
        }
    }
    z = bz * BZ + threadIdx.x;
    y = by * BY + threadIdx.y;
    return true;
}

// Waves per SIMD the velocity update is compiled for: FD orders 2 and 4 fit
// 80 VGPRs (6 waves instead of 4) without spilling; higher orders need more.
// (FP64 arithmetic needs twice the registers: no constraint there.)
constexpr int v3d_waves(int H, int rbytes) { return rbytes == 8 ? 2 : H <= 2 ? 6 : H <= 4 ? 4 : 2; }

template <typename R, typename S, int H>
__global__ __launch_bounds__(64 * BYV) __attribute__((amdgpu_waves_per_eu(v3d_waves(H, sizeof(R)), 8)))
void k_v3d(FDArgs<R> a, int nbz, unsigned nb, int yg) {
    using T = Store<R, S>;
    const Layout &g = a.g;
    int z, y, x;
    if (!tile3<BYV>(nbz, g.NX, g.NY, nb, yg, z, y, x)) return;
    if (z >= g.NZ || y >= g.NY) return;
    const int NAB = a.nab, NZ = g.NZ, NY = g.NY, NX = g.NX;
    const long sx = g.sx, sy = g.sy;
    const R *hc = a.hc;
    const long cell = g.idx(z, y, x);
    const unsigned o = (unsigned)(cell * sizeof(R));    // lane byte offset, coefficients
    const unsigned os = (unsigned)(cell * sizeof(S));   // lane byte offset, wavefields
    const S *szzp = (const S *)a.szz, *sxxp = (const S *)a.sxx, *sxzp = (const S *)a.sxz;
    const S *syzp = (const S *)a.syz, *syyp = (const S *)a.syy, *sxyp = (const S *)a.sxy;
    S *vxp = (S *)a.vx, *vyp = (S *)a.vy, *vzp = (S *)a.vz;
    R szz_z = DpS<H>(szzp, os, 1, hc);
    R sxx_x = DpS<H>(sxxp, os, sx, hc);
    R sxz_x = DmS<H>(sxzp, os, sx, hc);
    R sxz_z = DmS<H>(sxzp, os, 1, hc);
    R syz_y = DmS<H>(syzp, os, sy, hc);
    R syz_z = DmS<H>(syzp, os, 1, hc);
    R syy_y = DpS<H>(syyp, os, sy, hc);
    R sxy_x = DmS<H>(sxyp, os, sx, hc);
    R sxy_y = DmS<H>(sxyp, os, sy, hc);
    // CPML strips: every load first (psi and profiles, then the cell's own
    // values), the psi stores after
    const bool cpml = a.abs_type == 1;
    const bool pz = cpml && (z > NZ - NAB - 1 || (!a.freesurf && z < NAB));
    const bool py = cpml && (y < NAB || y > NY - NAB - 1);
    const bool px = cpml && (x < NAB || x > NX - NAB - 1);
    const int k = z < NAB ? z : z - NZ + 2 * NAB;
    const int ki = z < NAB ? k : 2 * NAB - k, kh = z < NAB ? k : 2 * NAB - 1 - k;
    const int j = y < NAB ? y : y - NY + 2 * NAB;
    const int ji = y < NAB ? j : 2 * NAB - j, jh = y < NAB ? j : 2 * NAB - 1 - j;
    const int ii = x < NAB ? x : x - NX + 2 * NAB;
    const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
    const long ipz = (long)x * NY * 2 * NAB + (long)y * 2 * NAB + k;
    const long ipy = (long)x * 2 * NAB * NZ + (long)j * NZ + z;
    const long ipx = (long)ii * NY * NZ + (long)y * NZ + z;
    Pml<R> z0, z1, z2, y0, y1, y2, x0, x1, x2;
    if (pz) {
        z0 = pml_load<kNTP>(a, a.psi_sxz_z + ipz, 2, false, ki);
        z1 = pml_load<kNTP>(a, a.psi_syz_z + ipz, 2, false, ki);
        z2 = pml_load<kNTP>(a, a.psi_szz_z + ipz, 2, true, kh);
    }
    if (py) {
        y0 = pml_load<kNTP>(a, a.psi_sxy_y + ipy, 1, false, ji);
        y1 = pml_load<kNTP>(a, a.psi_syy_y + ipy, 1, true, jh);
        y2 = pml_load<kNTP>(a, a.psi_syz_y + ipy, 1, false, ji);
    }
    if (px) {
        x0 = pml_load<kNTP>(a, a.psi_sxx_x + ipx, 0, true, iH);
        x1 = pml_load<kNTP>(a, a.psi_sxy_x + ipx, 0, false, iI);
        x2 = pml_load<kNTP>(a, a.psi_sxz_x + ipx, 0, false, iI);
    }
    auto ldv = [](const S *p, unsigned off) -> R {
        if constexpr (kNTV & 1) return T::ldn_(p, off);
        else return T::ld(p, off);
    };
    auto ldc = [](const R *p, unsigned off) -> R {
        if constexpr (kNTV & 2) return ldn(p, off);
        else return ldo(p, off);
    };
    const R vx0 = ldv(vxp, os), vy0 = ldv(vyp, os), vz0 = ldv(vzp, os);
    const R rip = ldc(a.rip, o), rjp = ldc(a.rjp, o), rkp = ldc(a.rkp, o);
    if (pz) {
        pml_apply<kNTP>(sxz_z, a.psi_sxz_z + ipz, z0);
        pml_apply<kNTP>(syz_z, a.psi_syz_z + ipz, z1);
        pml_apply<kNTP>(szz_z, a.psi_szz_z + ipz, z2);
    }
    if (py) {
        pml_apply<kNTP>(sxy_y, a.psi_sxy_y + ipy, y0);
        pml_apply<kNTP>(syy_y, a.psi_syy_y + ipy, y1);
        pml_apply<kNTP>(syz_y, a.psi_syz_y + ipy, y2);
    }
    if (px) {
        pml_apply<kNTP>(sxx_x, a.psi_sxx_x + ipx, x0);
        pml_apply<kNTP>(sxy_x, a.psi_sxy_x + ipx, x1);
        pml_apply<kNTP>(sxz_x, a.psi_sxz_x + ipx, x2);
    }
    R vx = vx0 + ((sxx_x + sxy_y + sxz_z) * rip);
    R vy = vy0 + ((syy_y + sxy_x + syz_z) * rjp);
    R vz = vz0 + ((szz_z + sxz_x + syz_y) * rkp);
    if (a.abs_type == 2) {
        const R *t = a.taper;
        if (!a.freesurf && z < NAB) { vx *= t[z]; vy *= t[z]; vz *= t[z]; }
        if (z > NZ - NAB - 1) { vx *= t[NZ - z - 1]; vy *= t[NZ - z - 1]; vz *= t[NZ - z - 1]; }
        if (y < NAB) { vx *= t[y]; vy *= t[y]; vz *= t[y]; }
        if (y > NY - NAB - 1) { vx *= t[NY - y - 1]; vy *= t[NY - y - 1]; vz *= t[NY - y - 1]; }
        if (x < NAB) { vx *= t[x]; vy *= t[x]; vz *= t[x]; }
        if (x > NX - NAB - 1) { vx *= t[NX - x - 1]; vy *= t[NX - x - 1]; vz *= t[NX - x - 1]; }
    }
    if constexpr (kNTV & 8) {
        T::stn_(vxp, os, vx);
        T::stn_(vyp, os, vy);
        T::stn_(vzp, os, vz);
    } else {
        T::st(vxp, os, vx);
        T::st(vyp, os, vy);
        T::st(vzp, os, vz);
    }
}

template <typename R, typename S, int H, int L, bool FS>
__global__ __launch_bounds__((64 * bys<R, S>())) void k_s3d(FDArgs<R> a, int nbz, unsigned nb, int yg) {
    using T = Store<R, S>;
    const Layout &g = a.g;
    int z, y, x;
    if (!tile3<bys<R, S>()>(nbz, g.NX, g.NY, nb, yg, z, y, x)) return;
    if (z >= g.NZ || y >= g.NY) return;
    const int NAB = a.nab, NZ = g.NZ, NY = g.NY, NX = g.NX;
    const long sx = g.sx, sy = g.sy, nf = g.nfield;
    const R *hc = a.hc;
    const long cell = g.idx(z, y, x);
    const unsigned o = (unsigned)(cell * sizeof(R));    // lane byte offset, coefficients
    const unsigned os = (unsigned)(cell * sizeof(S));   // lane byte offset, wavefields
    const S *vxp = (const S *)a.vx, *vyp = (const S *)a.vy, *vzp = (const S *)a.vz;
    S *sxxp = (S *)a.sxx, *syyp = (S *)a.syy, *szzp = (S *)a.szz;
    S *sxyp = (S *)a.sxy, *syzp = (S *)a.syz, *sxzp = (S *)a.sxz;
    S *rxxp = (S *)a.rxx, *ryyp = (S *)a.ryy, *rzzp = (S *)a.rzz;
    S *rxyp = (S *)a.rxy, *ryzp = (S *)a.ryz, *rxzp = (S *)a.rxz;
    R vzx = DpS<H>(vzp, os, sx, hc);
    R vzy = DpS<H>(vzp, os, sy, hc);
    R vzz = DmS<H>(vzp, os, 1, hc);
    R vyx = DpS<H>(vyp, os, sx, hc);
    R vyy = DmS<H>(vyp, os, sy, hc);
    R vyz = DpS<H>(vyp, os, 1, hc);
    R vxx = DmS<H>(vxp, os, sx, hc);
    R vxy = DpS<H>(vxp, os, sy, hc);
    R vxz = DpS<H>(vxp, os, 1, hc);
    const bool cpml = a.abs_type == 1;
    const bool pz = cpml && (z > NZ - NAB - 1 || (!a.freesurf && z < NAB));
    const bool py = cpml && (y < NAB || y > NY - NAB - 1);
    const bool px = cpml && (x < NAB || x > NX - NAB - 1);
    const int k = z < NAB ? z : z - NZ + 2 * NAB;
    const int ki = z < NAB ? k : 2 * NAB - k, kh = z < NAB ? k : 2 * NAB - 1 - k;
    const int j = y < NAB ? y : y - NY + 2 * NAB;
    const int ji = y < NAB ? j : 2 * NAB - j, jh = y < NAB ? j : 2 * NAB - 1 - j;
    const int ii = x < NAB ? x : x - NX + 2 * NAB;
    const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
    const long ipz = (long)x * NY * 2 * NAB + (long)y * 2 * NAB + k;
    const long ipy = (long)x * 2 * NAB * NZ + (long)j * NZ + z;
    const long ipx = (long)ii * NY * NZ + (long)y * NZ + z;
    Pml<R> z0, z1, z2, y0, y1, y2, x0, x1, x2;
    if (pz) {
        z0 = pml_load<kNTP>(a, a.psi_vx_z + ipz, 2, true, kh);
        z1 = pml_load<kNTP>(a, a.psi_vy_z + ipz, 2, true, kh);
        z2 = pml_load<kNTP>(a, a.psi_vz_z + ipz, 2, false, ki);
    }
    if (py) {
        y0 = pml_load<kNTP>(a, a.psi_vx_y + ipy, 1, true, jh);
        y1 = pml_load<kNTP>(a, a.psi_vy_y + ipy, 1, false, ji);
        y2 = pml_load<kNTP>(a, a.psi_vz_y + ipy, 1, true, jh);
    }
    if (px) {
        x0 = pml_load<kNTP>(a, a.psi_vx_x + ipx, 0, false, iI);
        x1 = pml_load<kNTP>(a, a.psi_vy_x + ipx, 0, true, iH);
        x2 = pml_load<kNTP>(a, a.psi_vz_x + ipx, 0, true, iH);
    }
    // the cell's own values and coefficients, before any store
    auto lds = [](const S *p, unsigned off) -> R {
        if constexpr (kNTS & 1) return T::ldn_(p, off);
        else return T::ld(p, off);
    };
    auto ldr = [](const S *p, unsigned off) -> R {
        if constexpr (kNTS & 4) return T::ldn_(p, off);
        else return T::ld(p, off);
    };
    auto ldc = [](const R *p, unsigned off) -> R {
        if constexpr (kNTS & 2) return ldn(p, off);
        else return ldo(p, off);
    };
    auto sts = [](S *p, unsigned off, R v) {
        if constexpr (kNTS & 8) T::stn_(p, off, v);
        else T::st(p, off, v);
    };
    R sxy = lds(sxyp, os), syz = lds(syzp, os), sxz = lds(sxzp, os);
    R sxx = lds(sxxp, os), syy = lds(syyp, os), szz = lds(szzp, os);
    R rxy[L > 0 ? L : 1], ryz[L > 0 ? L : 1], rxz[L > 0 ? L : 1];
    R rxx[L > 0 ? L : 1], ryy[L > 0 ? L : 1], rzz[L > 0 ? L : 1];
    R gg, f, d = 0, e = 0;
    R fipjp, fjpkp, fipkp, dipjp = 0, djpkp = 0, dipkp = 0;
    if constexpr (L == 0) {
        fipjp = ldc(a.muipjp, o);
        fjpkp = ldc(a.mujpkp, o);
        fipkp = ldc(a.muipkp, o);
        gg = ldc(a.M, o);
        f = (R)2 * ldc(a.mu, o);
    } else {
        gg = ldc(a.cg, o);
        f = ldc(a.cf, o);
        d = ldc(a.cd, o);
        e = ldc(a.ce, o);
        fipjp = ldc(a.fipjp, o);
        fjpkp = ldc(a.fjpkp, o);
        fipkp = ldc(a.fipkp, o);
        dipjp = ldc(a.dipjp, o);
        djpkp = ldc(a.djpkp, o);
        dipkp = ldc(a.dipkp, o);
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const long ln = l * nf;
            rxy[l] = ldr(rxyp + ln, os); ryz[l] = ldr(ryzp + ln, os); rxz[l] = ldr(rxzp + ln, os);
            rxx[l] = ldr(rxxp + ln, os); ryy[l] = ldr(ryyp + ln, os); rzz[l] = ldr(rzzp + ln, os);
        }
    }
    if (pz) {
        pml_apply<kNTP>(vxz, a.psi_vx_z + ipz, z0);
        pml_apply<kNTP>(vyz, a.psi_vy_z + ipz, z1);
        pml_apply<kNTP>(vzz, a.psi_vz_z + ipz, z2);
    }
    if (py) {
        pml_apply<kNTP>(vxy, a.psi_vx_y + ipy, y0);
        pml_apply<kNTP>(vyy, a.psi_vy_y + ipy, y1);
        pml_apply<kNTP>(vzy, a.psi_vz_y + ipy, y2);
    }
    if (px) {
        pml_apply<kNTP>(vxx, a.psi_vx_x + ipx, x0);
        pml_apply<kNTP>(vyx, a.psi_vy_x + ipx, x1);
        pml_apply<kNTP>(vzx, a.psi_vz_x + ipx, x2);
    }
    const R vxyyx = vxy + vyx;
    const R vyzzy = vyz + vzy;
    const R vxzzx = vxz + vzx;
    const R vxxyyzz = vxx + vyy + vzz;
    const R vyyzz = vyy + vzz;
    const R vxxzz = vxx + vzz;
    const R vxxyy = vxx + vyy;
    if constexpr (L == 0) {
        sxy += (fipjp * vxyyx);
        syz += (fjpkp * vyzzy);
        sxz += (fipkp * vxzzx);
        sxx += ((gg * vxxyyzz) - (f * vyyzz));
        syy += ((gg * vxxyyzz) - (f * vxxzz));
        szz += ((gg * vxxyyzz) - (f * vxxyy));
    } else {
        const R DT2 = a.DT2;
        R sumrxy = 0, sumryz = 0, sumrxz = 0, sumrxx = 0, sumryy = 0, sumrzz = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            sumrxy += rxy[l]; sumryz += ryz[l]; sumrxz += rxz[l];
            sumrxx += rxx[l]; sumryy += ryy[l]; sumrzz += rzz[l];
        }
        const R lsxy = (fipjp * vxyyx) + (DT2 * sumrxy);
        const R lsyz = (fjpkp * vyzzy) + (DT2 * sumryz);
        const R lsxz = (fipkp * vxzzx) + (DT2 * sumrxz);
        const R lsxx = ((gg * vxxyyzz) - (f * vyyzz)) + (DT2 * sumrxx);
        const R lsyy = ((gg * vxxyyzz) - (f * vxxzz)) + (DT2 * sumryy);
        const R lszz = ((gg * vxxyyzz) - (f * vxxyy)) + (DT2 * sumrzz);
        sumrxy = sumryz = sumrxz = sumrxx = sumryy = sumrzz = 0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const R b = a.bl[l], c = a.cl[l], et = a.eta[l];
            rxy[l] = b * (rxy[l] * c - et * (dipjp * vxyyx));
            ryz[l] = b * (ryz[l] * c - et * (djpkp * vyzzy));
            rxz[l] = b * (rxz[l] * c - et * (dipkp * vxzzx));
            rxx[l] = b * (rxx[l] * c - et * ((e * vxxyyzz) - (d * vyyzz)));
            ryy[l] = b * (ryy[l] * c - et * ((e * vxxyyzz) - (d * vxxzz)));
            rzz[l] = b * (rzz[l] * c - et * ((e * vxxyyzz) - (d * vxxyy)));
            sumrxy += rxy[l]; sumryz += ryz[l]; sumrxz += rxz[l];
            sumrxx += rxx[l]; sumryy += ryy[l]; sumrzz += rzz[l];
        }
        sxy += lsxy + (DT2 * sumrxy);
        syz += lsyz + (DT2 * sumryz);
        sxz += lsxz + (DT2 * sumrxz);
        sxx += lsxx + (DT2 * sumrxx);
        syy += lsyy + (DT2 * sumryy);
        szz += lszz + (DT2 * sumrzz);
    }
    if (a.abs_type == 2) {
        const R *t = a.taper;
        R *s6[6] = {&sxy, &syz, &sxz, &sxx, &syy, &szz};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            R &s = *s6[q];
            if (!a.freesurf && z < NAB) s *= t[z];
            if (z > NZ - NAB - 1) s *= t[NZ - z - 1];
            if (y < NAB) s *= t[y];
            if (y > NY - NAB - 1) s *= t[NY - y - 1];
            if (x < NAB) s *= t[x];
            if (x > NX - NAB - 1) s *= t[NX - x - 1];
        }
    }
    if constexpr (FS) {
        if (z == 0) {
            // image method at the surface row (src/surface3D.cl:84-252);
            // vxx, vyy (CPML-corrected) and vzz are those of this update
            szz = 0;
            if constexpr (L == 0) {
                const R h = -((gg - f) * (gg - f) * (vxx + vyy) / gg) - ((gg - f) * vzz);
                sxx += h;
                syy += h;
            } else {
                const R DT2 = a.DT2;
                const R h = -((gg - f) * (gg - f) * (vxx + vyy) / gg) - ((gg - f) * vzz);
                R sumx = 0, sumy = 0;
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    rzz[l] = 0;
                    sumx += rxx[l];
                    sumy += ryy[l];
                }
                sxx += (R)1 * (h - DT2 * sumx);
                syy += (R)1 * (h - DT2 * sumy);
                sumx = 0;
                sumy = 0;
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    const R hh = (R)((double)a.bs[l] *
                                     ((double)(d - e) * ((double)(f / gg) - 1.0) *
                                          (double)(vxx + vyy) -
                                      (double)((d - e) * vzz)));
                    rxx[l] += hh;
                    ryy[l] += hh;
                    sumx += rxx[l];
                    sumy += ryy[l];
                }
                sxx += (DT2 * sumx);
                syy += (DT2 * sumy);
            }
        }
    }
    sts(sxyp, os, sxy); sts(syzp, os, syz); sts(sxzp, os, sxz);
    sts(sxxp, os, sxx); sts(syyp, os, syy); sts(szzp, os, szz);
    if constexpr (L > 0) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const long ln = l * nf;
            sts(rxyp + ln, os, rxy[l]); sts(ryzp + ln, os, ryz[l]); sts(rxzp + ln, os, rxz[l]);
            sts(rxxp + ln, os, rxx[l]); sts(ryyp + ln, os, ryy[l]); sts(rzzp + ln, os, rzz[l]);
        }
    }
    if constexpr (FS) {
        // mirror images above the surface: szz(-m) = -szz(m),
        // sxz(-m) = -sxz(m-1), syz(-m) = -syz(m-1), m = 1..H
        if (z >= 1 && z <= H) T::st(szzp, os - 2 * z * (unsigned)sizeof(S), -szz);
        if (z <= H - 1) {
            T::st(sxzp, os - (2 * z + 1) * (unsigned)sizeof(S), -sxz);
            T::st(syzp, os - (2 * z + 1) * (unsigned)sizeof(S), -syz);
        }
    }
}

template <int BY>
void grid3(const Layout &g, int &nbz, unsigned &nb) {
    nbz = (g.NZ + BZ - 1) / BZ;
    const int nby = (g.NY + BY - 1) / BY;
    nb = (unsigned)nbz * g.NX * nby;
}

// Slabs per y group (tile3): the largest of want, want/2, ... that divides
// the slab count; want comes from SEIS_YGROUP_V / SEIS_YGROUP_S when set
// (tuning), else from the defaults below.  Stress update: groups of 4 slabs
// (32 rows) measured 0.661 ms against 0.689 ms ungrouped (viscoelastic
// 256^3, FD4, CPML, free surface); the velocity update and the elastic
// stress update measured the same with any grouping.
#ifndef SEIS_YG_V
#define SEIS_YG_V 1
#endif
#ifndef SEIS_YG_S
#define SEIS_YG_S 4
#endif
inline int env_or(const char *name, int def) {
    const char *v = getenv(name);
    return v && *v ? atoi(v) : def;
}
template <int BY>
int ygroup(const Layout &g, int want) {
    const int nby = (g.NY + BY - 1) / BY;
    int G = want < 1 ? 1 : want;
    while (G > 1 && nby % G) G /= 2;
    return G;
}

template <typename R, typename S, int H>
void run_v3d(const FDArgs<R> &a, hipStream_t st) {
    int nbz;
    unsigned nb;
    grid3<BYV>(a.g, nbz, nb);
    static const int want = env_or("SEIS_YGROUP_V", SEIS_YG_V);
    const int yg = ygroup<BYV>(a.g, want);
    hipLaunchKernelGGL((k_v3d<R, S, H>), dim3(((nb + 7) / 8) * 8), dim3(BZ, BYV), 0, st, a, nbz,
                       nb, yg);
}

template <typename R, typename S, int H, int L>
void run_s3d(const FDArgs<R> &a, hipStream_t st) {
    int nbz;
    unsigned nb;
    grid3<bys<R, S>()>(a.g, nbz, nb);
    static const int want = env_or("SEIS_YGROUP_S", SEIS_YG_S);
    const int yg = ygroup<bys<R, S>()>(a.g, want);
    const dim3 grid(((nb + 7) / 8) * 8), blk(BZ, bys<R, S>());
    if (a.freesurf && a.fuse_surface)
        hipLaunchKernelGGL((k_s3d<R, S, H, L, true>), grid, blk, 0, st, a, nbz, nb, yg);
    else
        hipLaunchKernelGGL((k_s3d<R, S, H, L, false>), grid, blk, 0, st, a, nbz, nb, yg);
}

template <typename R, typename S, int H>
void run_s3d_L(const FDArgs<R> &a, hipStream_t st) {
    switch (a.L) {
        case 0: run_s3d<R, S, H, 0>(a, st); break;
        case 1: run_s3d<R, S, H, 1>(a, st); break;
        case 2: run_s3d<R, S, H, 2>(a, st); break;
        case 3: run_s3d<R, S, H, 3>(a, st); break;
        case 4: run_s3d<R, S, H, 4>(a, st); break;
        default: break;
    }
}

}  // namespace

template <typename R, typename S>
void launch_v3d(const FDArgs<R> &a, void *stream) {
    SEIS_SWITCH_H(a.g.H, (run_v3d<R, S, HH>(a, (hipStream_t)stream)));
}
template <typename R, typename S>
void launch_s3d(const FDArgs<R> &a, void *stream) {
    SEIS_SWITCH_H(a.g.H, (run_s3d_L<R, S, HH>(a, (hipStream_t)stream)));
}

template void launch_v3d<float, float>(const FDArgs<float> &, void *);
template void launch_s3d<float, float>(const FDArgs<float> &, void *);
template void launch_v3d<float, __half>(const FDArgs<float> &, void *);
template void launch_s3d<float, __half>(const FDArgs<float> &, void *);
template void launch_v3d<double, double>(const FDArgs<double> &, void *);
template void launch_s3d<double, double>(const FDArgs<double> &, void *);

}  // namespace seis
