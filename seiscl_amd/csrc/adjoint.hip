// Adjoint step kernels of the gradient by backpropagation (elastic, 2D/3D).
//
// Each kernel does, for the cell its lane owns and in the reference's order
// (src/update_adjs2D.cl, src/update_adjv2D.cl, src/update_adjs3D.cl,
// src/update_adjv3D.cl):
//   1. backpropagate the forward field: subtract this step's increment,
//      computed without CPML from the other forward field;
//   2. overwrite it with the saved boundary value when the cell lies on the
//      boundary shell (inject_ind);
//   3. update the adjoint field exactly as the forward kernels update theirs
//      (same stencils, CPML on the adjoint run's own memory, taper);
//   4. accumulate the gradient, in double, from the backpropagated forward
//      field and the adjoint increment.
// Every access of a lane is to its own cell, except stencil reads of the
// other field, which no kernel writes, so the steps need no barriers.
//
// Also here: boundary saving (savebnd) and adjoint-source injection
// (residuals, src/automatic_kernels.c:631-836).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fd_common.hpp"
#include "seis_internal.hpp"

namespace seis {
namespace {

constexpr int ABY = 4;   // rows of y (3D) or x (2D) per block

// 3D kernels: non-temporal loads and stores (the nt bit) for the streams a
// lane touches once (its own cell's fields, memory variables, coefficients,
// forward history and gradient accumulators), so that the stencil
// neighbours read by co-resident blocks stay in L2.
#ifndef SEIS_NT_ADJ
#define SEIS_NT_ADJ 1
#endif
template <typename T>
__device__ __forceinline__ T own_ld(const T *p) {
    if constexpr (SEIS_NT_ADJ != 0) return __builtin_nontemporal_load(p);
    else return *p;
}
template <typename T>
__device__ __forceinline__ void own_st(T *p, T v) {
    if constexpr (SEIS_NT_ADJ != 0) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <typename T>
__device__ __forceinline__ void own_add(T *p, T v) { own_st(p, own_ld(p) + v); }

// CPML correction of one adjoint derivative on the near/far strip of an
// axis, with the profile index of the reference kernels (full grid: k or
// 2*NAB-k; half grid: k or 2*NAB-1-k).
template <typename R>
__device__ __forceinline__ void pml_full(const FDArgs<R> &a, int axis, R &d, R *psi, int ki) {
    pml(d, psi, cp(a, axis, CB, ki), cp(a, axis, CA, ki), cp(a, axis, CK, ki));
}
template <typename R>
__device__ __forceinline__ void pml_half(const FDArgs<R> &a, int axis, R &d, R *psi, int kh) {
    pml(d, psi, cp(a, axis, CBH, kh), cp(a, axis, CAH, kh), cp(a, axis, CKH, kh));
}

// taper factor product of the cell (abs_type 2), applied factor by factor
template <typename R>
__device__ __forceinline__ void taper_apply(const FDArgs<R> &a, int z, int y, int x, R *v, int n) {
    const R *t = a.taper;
    const int NAB = a.nab, NZ = a.g.NZ, NY = a.g.NY, NX = a.g.NX;
    if (!a.freesurf && z < NAB)
        for (int q = 0; q < n; ++q) v[q] *= t[z];
    if (z > NZ - NAB - 1)
        for (int q = 0; q < n; ++q) v[q] *= t[NZ - z - 1];
    if (a.g.nd == 3) {
        if (y < NAB)
            for (int q = 0; q < n; ++q) v[q] *= t[y];
        if (y > NY - NAB - 1)
            for (int q = 0; q < n; ++q) v[q] *= t[NY - y - 1];
    }
    if (x < NAB)
        for (int q = 0; q < n; ++q) v[q] *= t[x];
    if (x > NX - NAB - 1)
        for (int q = 0; q < n; ++q) v[q] *= t[NX - x - 1];
}

// lane -> cell, z fastest; 2D uses y = 0
__device__ __forceinline__ bool adj_cell(const Layout &g, int nbz, unsigned nb, int &z, int &y,
                                         int &x) {
    unsigned tile;
    xcd_tile(nb, tile);
    if (tile >= nb) return false;
    const int bz = tile % nbz;
    const unsigned r = tile / nbz;
    z = bz * BZ + threadIdx.x;
    if (g.nd == 3) {
        const int nby = (g.NY + ABY - 1) / ABY;
        y = (r % nby) * ABY + threadIdx.y;
        x = r / nby;
        return z < g.NZ && y < g.NY && x < g.NX;
    }
    y = 0;
    x = r * ABY + threadIdx.y;
    return z < g.NZ && x < g.NX;
}

// The boundary shell of this step: A.bnd, or, in replayed step graphs,
// A.bnd + (*A.tb + A.bofs) steps (a device step counter plus the step's
// offset within the graph).
template <typename R>
__device__ __forceinline__ const R *adj_bnd(const AdjArgs<R> &A) {
    return A.tb ? A.bnd + ((long)*A.tb + A.bofs) * A.bper : A.bnd;
}

// ---------------------------------------------------------------------------
// 2D
// ---------------------------------------------------------------------------
// M and mu gradient increments of one cell from the forward stresses s*
// and the adjoint stress increments l* (src/update_adjs2D.cl:441-468), plus
// the approximate Hessian (Hout, :453-456) when A.hM is set; restype 1
// accumulates the RTM image of M instead.  Every increment is taken to
// physical units in double (factor 1 unless the fields are stored scaled).
template <typename R>
__device__ __forceinline__ void grad_mmu_2d(const AdjArgs<R> &A, long i, R sxx, R szz, R sxz,
                                            R lsxx, R lszz, R lsxz, R lM, R lmu) {
    if (A.rtm) {
        const R dM = (sxx + szz) * (lsxx + lszz);
        A.gM[i] += (double)-dM * A.gscale_s;
        if (A.hM) A.hM[i] += (double)((sxx + szz) * (sxx + szz)) * A.hscale;
        return;
    }
    const R c1 = 1.0 / ((2.0 * lM - 2.0 * lmu) * (2.0 * lM - 2.0 * lmu));
    const R c3 = 1.0 / (lmu * lmu);
    const R c5 = 0.25 * c3;
    const R dM = c1 * (sxx + szz) * (lsxx + lszz);
    if (A.gsh[0]) {
        // the shear term per staggered node, chained on the host
        const R dmu = dM - c5 * ((sxx - szz) * (lsxx - lszz));
        A.gsh[0][i] += (double)(sxz * lsxz) * A.gscale_s;
        A.gmu[i] += (double)dmu * A.gscale_s;
    } else {
        const R dmu = -c3 * (sxz * lsxz) + dM - c5 * ((sxx - szz) * (lsxx - lszz));
        A.gmu[i] += (double)dmu * A.gscale_s;
    }
    A.gM[i] += (double)-dM * A.gscale_s;
    if (A.hM) {
        const R dMH = c1 * (sxx + szz) * (sxx + szz);
        const R hmu = c3 * sxz * sxz - (A.hfix ? dMH : dM) + c5 * (sxx - szz) * (sxx - szz);
        A.hM[i] += (double)dMH * A.hscale;
        A.hmu[i] += (double)hmu * A.hscale;
    }
}

// The same in 3D (src/update_adjs3D.cl:678-710).  The reference's mu
// Hessian expression does not compile (one parenthesis too many); here c5
// weighs the sum of the three squared deviatoric stresses, and the M term is
// the Hessian's own, as the 2D form with hfix.
template <typename R>
__device__ __forceinline__ void grad_mmu_3d(const AdjArgs<R> &A, long i, R sxx, R syy, R szz,
                                            R sxy, R syz, R sxz, R lsxx, R lsyy, R lszz,
                                            R lsxy, R lsyz, R lsxz, R lM, R lmu) {
    if (A.rtm) {
        const R dM = (sxx + syy + szz) * (lsxx + lsyy + lszz);
        own_add(A.gM + i, (double)-dM * A.gscale_s);
        if (A.hM) own_add(A.hM + i, (double)((sxx + syy + szz) * (sxx + syy + szz)) * A.hscale);
        return;
    }
    const R c1 = 1.0 / (3.0 * lM - 4.0 * lmu) / (3.0 * lM - 4.0 * lmu);
    const R c3 = 1.0 / lmu / lmu;
    const R c5 = 1.0 / 6.0 * c3;
    const R dM = c1 * (sxx + syy + szz) * (lsxx + lsyy + lszz);
    // evaluated in double from the 4.0 / 3 term on, as the reference's
    // unsuffixed literals make it
    const double dev = c5 * (lsxx * (2.0 * sxx - syy - szz) + lsyy * (2.0 * syy - sxx - szz) +
                             lszz * (2.0 * szz - sxx - syy));
    if (A.gsh[0]) {
        // the shear terms per staggered node, chained on the host
        own_add(A.gsh[0] + i, (double)(sxz * lsxz) * A.gscale_s);
        own_add(A.gsh[1] + i, (double)(sxy * lsxy) * A.gscale_s);
        own_add(A.gsh[2] + i, (double)(syz * lsyz) * A.gscale_s);
        own_add(A.gmu + i, (4.0 / 3 * dM - dev) * A.gscale_s);
    } else {
        const double dmu = -c3 * (sxz * lsxz + sxy * lsxy + syz * lsyz) + 4.0 / 3 * dM - dev;
        own_add(A.gmu + i, dmu * A.gscale_s);
    }
    own_add(A.gM + i, (double)-dM * A.gscale_s);
    if (A.hM) {
        const R dMH = c1 * (sxx + syy + szz) * (sxx + syy + szz);
        const double hmu = c3 * (sxz * sxz + sxy * sxy + syz * syz) - 4.0 / 3 * dMH +
                           c5 * ((2.0 * sxx - syy - szz) * (2.0 * sxx - syy - szz) +
                                 (2.0 * syy - sxx - szz) * (2.0 * syy - sxx - szz) +
                                 (2.0 * szz - sxx - syy) * (2.0 * szz - sxx - syy));
        own_add(A.hM + i, (double)dMH * A.hscale);
        own_add(A.hmu + i, hmu * A.hscale);
    }
}

// tau gradients of the GSLS stress update (no counterpart in the
// reference's time-domain kernels, which declare gradtaup/gradtaus and never
// write them, src/update_adjs3D.cl:66; its DFT path forms them in
// src/calc_grad.c:111-152).  A normal-stress update is
//   inc_ii = g th - f (th - v_ii) + Rm_ii,  g = M(1+L tp), f = 2mu(1+L ts),
// Rm_ii = DT2 sum_l (r_ii,l old + new): the whole memory term of the step,
// the memory variables forced by e th - d (th - v_ii), e = M tp/DT,
// d = 2 mu ts/DT, through the same relaxation filters, so Rm_ii = e F -
// d F_ii and sum_ii Rm_ii = (ND e - (ND-1) d) F.  The change of the update
// with tp (at fixed M) is M L th + (M/DT) F, with ts (at fixed mu)
// -2mu L (th - v_ii) - (e F - Rm_ii)/ts, and for a shear component (pair
// mu_ij(1+L ts_ij), mu_ij ts_ij/DT) mu_ij L w_ij + Rm_ij/ts_ij.  Holding the
// wave speeds through the normalisation M = M_f0/(1 + s0 tp) adds
// -s0/(1+s0 tp) times the part linear in M (likewise for mu).  M, mu, tp, ts
// are recovered from g, f, d, e.  Cells with ts = 0 drop the memory term of
// ts.  The gradient is minus the forward strain (the forward stresses
// through the compliance of the gradient moduli, as for M and mu) times
// these changes.
template <typename R>
__device__ __forceinline__ void tau_parts_normal(R theta, const R *vii, const R *Rm, int nd, R g,
                                                 R f, R d, R e, R L, R DT, R s0, R &Xp, R *Xs) {
    const R MR = g - L * e * DT;
    const R tp = e * DT / MR;
    const R tmu = f - L * d * DT;
    const R ts = tmu != R(0) ? d * DT / tmu : R(0);
    const R kp = s0 / (R(1) + s0 * tp);
    const R ks = s0 / (R(1) + s0 * ts);
    const R den = nd == 3 ? R(3) * e - R(2) * d : R(2) * e - d;
    const R trm = nd == 3 ? (Rm[0] + Rm[1]) + Rm[2] : Rm[0] + Rm[1];
    const R F = den != R(0) ? trm / den : R(0);
    Xp = (MR * L - kp * g) * theta + (MR / DT - kp * e) * F;
    for (int k = 0; k < nd; ++k) {
        const R dev = theta - vii[k];
        const R q = e * F - Rm[k];
        const R Sk = -(f * dev) - q;
        Xs[k] = -(tmu * L * dev) - (ts > R(0) ? q / ts : R(0)) - ks * Sk;
    }
}

template <typename R>
__device__ __forceinline__ R tau_part_shear(R w, R Rm, R fij, R dij, R inc, R L, R DT, R s0) {
    const R mua = fij - L * dij * DT;
    const R tsa = mua != R(0) ? dij * DT / mua : R(0);
    const R ksa = s0 / (R(1) + s0 * tsa);
    return mua * L * w + (tsa > R(0) ? Rm / tsa : R(0)) - ksa * inc;
}

template <typename R>
__device__ __forceinline__ void grad_tau_2d(const AdjArgs<R> &A, long i, R sxx, R szz, R sxz,
                                            R lM, R lmu, R Xp, R Xxx, R Xzz, R Xxz) {
    const R lam = lM - R(2) * lmu;
    const R th = (sxx + szz) / (R(2) * lM - R(2) * lmu);
    A.gtp[i] += (double)(-(th * Xp)) * A.gscale_s;
    if (lmu == R(0)) return;
    const R exx = (sxx - lam * th) / (R(2) * lmu);
    const R ezz = (szz - lam * th) / (R(2) * lmu);
    const R gxz = sxz / lmu;
    A.gts[i] += (double)(-((exx * Xxx + ezz * Xzz) + gxz * Xxz)) * A.gscale_s;
}

template <typename R>
__device__ __forceinline__ void grad_tau_3d(const AdjArgs<R> &A, long i, R sxx, R syy, R szz,
                                            R sxy, R syz, R sxz, R lM, R lmu, R Xp,
                                            const R *X) {
    const R lam = lM - R(2) * lmu;
    const R th = ((sxx + syy) + szz) / (R(3) * lM - R(4) * lmu);
    own_add(A.gtp + i, (double)(-(th * Xp)) * A.gscale_s);
    if (lmu == R(0)) return;
    const R exx = (sxx - lam * th) / (R(2) * lmu);
    const R eyy = (syy - lam * th) / (R(2) * lmu);
    const R ezz = (szz - lam * th) / (R(2) * lmu);
    const R n3 = (exx * X[0] + eyy * X[1]) + ezz * X[2];
    const R s3 = ((sxy / lmu) * X[3] + (syz / lmu) * X[4]) + (sxz / lmu) * X[5];
    own_add(A.gts + i, (double)(-(n3 + s3)) * A.gscale_s);
}

template <typename R, int H>
__global__ __launch_bounds__(256) void k_adj_s2d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &f = A.f, &a = A.r;
    const Layout &g = f.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const R *hc = f.hc;
    const R fipkp = f.muipkp[i], lmu = f.mu[i], lM = f.M[i];
    const R ff = 2.0 * lmu, gg = lM;
    // 1-2. backpropagate the forward stresses, re-inject the shell
    R sxz = f.sxz[i], sxx = f.sxx[i], szz = f.szz[i];
    {
        const R vxx = Dm<H>(f.vx, i, sx, hc), vxz = Dp<H>(f.vx, i, 1, hc);
        const R vzz = Dm<H>(f.vz, i, 1, hc), vzx = Dp<H>(f.vz, i, sx, hc);
        sxz -= (fipkp * (vxz + vzx));
        sxx -= (gg * (vxx + vzz)) - (ff * vzz);
        szz -= (gg * (vxx + vzz)) - (ff * vxx);
        const long m = A.sh.inject(z + H, 0, x + H);
        if (m >= 0) {
            const long n = A.sh.n;
            const R *bnd = adj_bnd(A);
            sxx = bnd[2 * n + m];
            szz = bnd[3 * n + m];
            sxz = bnd[4 * n + m];
        }
        f.sxx[i] = sxx;
        f.szz[i] = szz;
        f.sxz[i] = sxz;
    }
    // 3. adjoint stress update
    R vxx = Dm<H>(a.vx, i, sx, hc), vxz = Dp<H>(a.vx, i, 1, hc);
    R vzz = Dm<H>(a.vz, i, 1, hc), vzx = Dp<H>(a.vz, i, sx, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const long ip = (long)x * 2 * NAB + k;
            pml_half(a, 2, vxz, a.psi_vx_z + ip, bot ? 2 * NAB - 1 - k : k);
            pml_full(a, 2, vzz, a.psi_vz_z + ip, bot ? 2 * NAB - k : k);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const long ip = (long)ii * NZ + z;
            pml_full(a, 0, vxx, a.psi_vx_x + ip, x < NAB ? ii : 2 * NAB - ii);
            pml_half(a, 0, vzx, a.psi_vz_x + ip, x < NAB ? ii : 2 * NAB - 1 - ii);
        }
    }
    const R lsxz = (fipkp * (vxz + vzx));
    const R lsxx = (gg * (vxx + vzz)) - (ff * vzz);
    const R lszz = (gg * (vxx + vzz)) - (ff * vxx);
    R s3[3] = {a.sxx[i] + lsxx, a.szz[i] + lszz, a.sxz[i] + lsxz};
    if (a.abs_type == 2) taper_apply(a, z, 0, x, s3, 3);
    a.sxx[i] = s3[0];
    a.szz[i] = s3[1];
    a.sxz[i] = s3[2];
    // 4. M and mu gradients (src/update_adjs2D.cl:441-468)
    grad_mmu_2d(A, i, sxx, szz, sxz, lsxx, lszz, lsxz, lM, lmu);
}

template <typename R, int H>
__global__ __launch_bounds__(256) void k_adj_v2d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &f = A.f, &a = A.r;
    const Layout &g = f.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const R *hc = f.hc;
    const R rip = f.rip[i], rkp = f.rkp[i];
    R vx = f.vx[i], vz = f.vz[i];
    {
        const R sxx_x = Dp<H>(f.sxx, i, sx, hc), szz_z = Dp<H>(f.szz, i, 1, hc);
        const R sxz_z = Dm<H>(f.sxz, i, 1, hc), sxz_x = Dm<H>(f.sxz, i, sx, hc);
        vx -= ((sxx_x + sxz_z) * rip);
        vz -= ((szz_z + sxz_x) * rkp);
        const long m = A.sh.inject(z + H, 0, x + H);
        if (m >= 0) {
            const R *bnd = adj_bnd(A);
            vx = bnd[m];
            vz = bnd[A.sh.n + m];
        }
        f.vx[i] = vx;
        f.vz[i] = vz;
    }
    R sxx_x = Dp<H>(a.sxx, i, sx, hc), szz_z = Dp<H>(a.szz, i, 1, hc);
    R sxz_z = Dm<H>(a.sxz, i, 1, hc), sxz_x = Dm<H>(a.sxz, i, sx, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const long ip = (long)x * 2 * NAB + k;
            pml_full(a, 2, sxz_z, a.psi_sxz_z + ip, bot ? 2 * NAB - k : k);
            pml_half(a, 2, szz_z, a.psi_szz_z + ip, bot ? 2 * NAB - 1 - k : k);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const long ip = (long)ii * NZ + z;
            pml_half(a, 0, sxx_x, a.psi_sxx_x + ip, x < NAB ? ii : 2 * NAB - 1 - ii);
            pml_full(a, 0, sxz_x, a.psi_sxz_x + ip, x < NAB ? ii : 2 * NAB - ii);
        }
    }
    const R lvx = ((sxx_x + sxz_z) * rip);
    const R lvz = ((szz_z + sxz_x) * rkp);
    R v2[2] = {a.vx[i] + lvx, a.vz[i] + lvz};
    if (a.abs_type == 2) taper_apply(a, z, 0, x, v2, 2);
    a.vx[i] = v2[0];
    a.vz[i] = v2[1];
    // density gradient (src/update_adjv2D.cl:326)
    if (A.grz) {
        A.grho[i] += (double)(-vx * lvx);
        A.grz[i] += (double)(-(vz * lvz));
    } else {
        const R d = -vx * lvx - vz * lvz;
        A.grho[i] += d;
    }
}

// ---------------------------------------------------------------------------
// 3D
// ---------------------------------------------------------------------------
template <typename R, int H>
__global__ __launch_bounds__(256) void k_adj_s3d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &f = A.f, &a = A.r;
    const Layout &g = f.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, y, x), sx = g.sx, sy = g.sy;
    const R *hc = f.hc;
    const R fipjp = own_ld(f.muipjp + i), fjpkp = own_ld(f.mujpkp + i), fipkp = own_ld(f.muipkp + i);
    const R lM = own_ld(f.M + i), lmu = own_ld(f.mu + i);
    const R gg = lM, ff = 2.0 * lmu;
    R sxx = own_ld(f.sxx + i), syy = own_ld(f.syy + i), szz = own_ld(f.szz + i);
    R sxy = own_ld(f.sxy + i), syz = own_ld(f.syz + i), sxz = own_ld(f.sxz + i);
    {
        const R vxx = Dm<H>(f.vx, i, sx, hc), vxy = Dp<H>(f.vx, i, sy, hc), vxz = Dp<H>(f.vx, i, 1, hc);
        const R vyx = Dp<H>(f.vy, i, sx, hc), vyy = Dm<H>(f.vy, i, sy, hc), vyz = Dp<H>(f.vy, i, 1, hc);
        const R vzx = Dp<H>(f.vz, i, sx, hc), vzy = Dp<H>(f.vz, i, sy, hc), vzz = Dm<H>(f.vz, i, 1, hc);
        const R vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
        const R vxxyyzz = vxx + vyy + vzz, vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
        sxy -= (fipjp * vxyyx);
        syz -= (fjpkp * vyzzy);
        sxz -= (fipkp * vxzzx);
        sxx -= ((gg * vxxyyzz) - (ff * vyyzz));
        syy -= ((gg * vxxyyzz) - (ff * vxxzz));
        szz -= ((gg * vxxyyzz) - (ff * vxxyy));
        const long m = A.sh.inject(z + H, y + H, x + H);
        if (m >= 0) {
            const long n = A.sh.n;
            const R *bnd = adj_bnd(A);
            sxx = bnd[3 * n + m];
            syy = bnd[4 * n + m];
            szz = bnd[5 * n + m];
            sxy = bnd[6 * n + m];
            syz = bnd[7 * n + m];
            sxz = bnd[8 * n + m];
        }
        own_st(f.sxx + i, sxx);
        own_st(f.syy + i, syy);
        own_st(f.szz + i, szz);
        own_st(f.sxy + i, sxy);
        own_st(f.syz + i, syz);
        own_st(f.sxz + i, sxz);
    }
    R vxx = Dm<H>(a.vx, i, sx, hc), vxy = Dp<H>(a.vx, i, sy, hc), vxz = Dp<H>(a.vx, i, 1, hc);
    R vyx = Dp<H>(a.vy, i, sx, hc), vyy = Dm<H>(a.vy, i, sy, hc), vyz = Dp<H>(a.vy, i, 1, hc);
    R vzx = Dp<H>(a.vz, i, sx, hc), vzy = Dp<H>(a.vz, i, sy, hc), vzz = Dm<H>(a.vz, i, 1, hc);
    const int NAB = a.nab, NZ = g.NZ, NY = g.NY, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int ki = bot ? 2 * NAB - k : k, kh = bot ? 2 * NAB - 1 - k : k;
            const long ip = (long)x * NY * 2 * NAB + (long)y * 2 * NAB + k;
            pml_half(a, 2, vxz, a.psi_vx_z + ip, kh);
            pml_half(a, 2, vyz, a.psi_vy_z + ip, kh);
            pml_full(a, 2, vzz, a.psi_vz_z + ip, ki);
        }
        if (y < NAB || y > NY - NAB - 1) {
            const int j = y < NAB ? y : y - NY + 2 * NAB;
            const int ji = y < NAB ? j : 2 * NAB - j, jh = y < NAB ? j : 2 * NAB - 1 - j;
            const long ip = (long)x * 2 * NAB * NZ + (long)j * NZ + z;
            pml_half(a, 1, vxy, a.psi_vx_y + ip, jh);
            pml_full(a, 1, vyy, a.psi_vy_y + ip, ji);
            pml_half(a, 1, vzy, a.psi_vz_y + ip, jh);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            const long ip = (long)ii * NY * NZ + (long)y * NZ + z;
            pml_full(a, 0, vxx, a.psi_vx_x + ip, iI);
            pml_half(a, 0, vyx, a.psi_vy_x + ip, iH);
            pml_half(a, 0, vzx, a.psi_vz_x + ip, iH);
        }
    }
    const R vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
    const R vxxyyzz = vxx + vyy + vzz, vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
    const R lsxy = (fipjp * vxyyx), lsyz = (fjpkp * vyzzy), lsxz = (fipkp * vxzzx);
    const R lsxx = ((gg * vxxyyzz) - (ff * vyyzz));
    const R lsyy = ((gg * vxxyyzz) - (ff * vxxzz));
    const R lszz = ((gg * vxxyyzz) - (ff * vxxyy));
    R s6[6] = {own_ld(a.sxy + i) + lsxy, own_ld(a.syz + i) + lsyz, own_ld(a.sxz + i) + lsxz,
               own_ld(a.sxx + i) + lsxx, own_ld(a.syy + i) + lsyy, own_ld(a.szz + i) + lszz};
    if (a.abs_type == 2) taper_apply(a, z, y, x, s6, 6);
    own_st(a.sxy + i, s6[0]);
    own_st(a.syz + i, s6[1]);
    own_st(a.sxz + i, s6[2]);
    own_st(a.sxx + i, s6[3]);
    own_st(a.syy + i, s6[4]);
    own_st(a.szz + i, s6[5]);
    // M and mu gradients (src/update_adjs3D.cl:678-710)
    grad_mmu_3d(A, i, sxx, syy, szz, sxy, syz, sxz, lsxx, lsyy, lszz, lsxy, lsyz, lsxz, lM, lmu);
}

template <typename R, int H>
__global__ __launch_bounds__(256) void k_adj_v3d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &f = A.f, &a = A.r;
    const Layout &g = f.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, y, x), sx = g.sx, sy = g.sy;
    const R *hc = f.hc;
    const R rip = own_ld(f.rip + i), rjp = own_ld(f.rjp + i), rkp = own_ld(f.rkp + i);
    R vx = own_ld(f.vx + i), vy = own_ld(f.vy + i), vz = own_ld(f.vz + i);
    {
        const R szz_z = Dp<H>(f.szz, i, 1, hc), sxx_x = Dp<H>(f.sxx, i, sx, hc);
        const R sxz_x = Dm<H>(f.sxz, i, sx, hc), sxz_z = Dm<H>(f.sxz, i, 1, hc);
        const R syz_y = Dm<H>(f.syz, i, sy, hc), syz_z = Dm<H>(f.syz, i, 1, hc);
        const R syy_y = Dp<H>(f.syy, i, sy, hc), sxy_x = Dm<H>(f.sxy, i, sx, hc);
        const R sxy_y = Dm<H>(f.sxy, i, sy, hc);
        vx -= ((sxx_x + sxy_y + sxz_z) * rip);
        vy -= ((syy_y + sxy_x + syz_z) * rjp);
        vz -= ((szz_z + sxz_x + syz_y) * rkp);
        const long m = A.sh.inject(z + H, y + H, x + H);
        if (m >= 0) {
            const long n = A.sh.n;
            const R *bnd = adj_bnd(A);
            vx = bnd[m];
            vy = bnd[n + m];
            vz = bnd[2 * n + m];
        }
        own_st(f.vx + i, vx);
        own_st(f.vy + i, vy);
        own_st(f.vz + i, vz);
    }
    R szz_z = Dp<H>(a.szz, i, 1, hc), sxx_x = Dp<H>(a.sxx, i, sx, hc);
    R sxz_x = Dm<H>(a.sxz, i, sx, hc), sxz_z = Dm<H>(a.sxz, i, 1, hc);
    R syz_y = Dm<H>(a.syz, i, sy, hc), syz_z = Dm<H>(a.syz, i, 1, hc);
    R syy_y = Dp<H>(a.syy, i, sy, hc), sxy_x = Dm<H>(a.sxy, i, sx, hc);
    R sxy_y = Dm<H>(a.sxy, i, sy, hc);
    const int NAB = a.nab, NZ = g.NZ, NY = g.NY, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int ki = bot ? 2 * NAB - k : k, kh = bot ? 2 * NAB - 1 - k : k;
            const long ip = (long)x * NY * 2 * NAB + (long)y * 2 * NAB + k;
            pml_full(a, 2, sxz_z, a.psi_sxz_z + ip, ki);
            pml_full(a, 2, syz_z, a.psi_syz_z + ip, ki);
            pml_half(a, 2, szz_z, a.psi_szz_z + ip, kh);
        }
        if (y < NAB || y > NY - NAB - 1) {
            const int j = y < NAB ? y : y - NY + 2 * NAB;
            const int ji = y < NAB ? j : 2 * NAB - j, jh = y < NAB ? j : 2 * NAB - 1 - j;
            const long ip = (long)x * 2 * NAB * NZ + (long)j * NZ + z;
            pml_full(a, 1, sxy_y, a.psi_sxy_y + ip, ji);
            pml_half(a, 1, syy_y, a.psi_syy_y + ip, jh);
            pml_full(a, 1, syz_y, a.psi_syz_y + ip, ji);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            const long ip = (long)ii * NY * NZ + (long)y * NZ + z;
            pml_half(a, 0, sxx_x, a.psi_sxx_x + ip, iH);
            pml_full(a, 0, sxy_x, a.psi_sxy_x + ip, iI);
            pml_full(a, 0, sxz_x, a.psi_sxz_x + ip, iI);
        }
    }
    const R lvx = ((sxx_x + sxy_y + sxz_z) * rip);
    const R lvy = ((syy_y + sxy_x + syz_z) * rjp);
    const R lvz = ((szz_z + sxz_x + syz_y) * rkp);
    R v3[3] = {own_ld(a.vx + i) + lvx, own_ld(a.vy + i) + lvy, own_ld(a.vz + i) + lvz};
    if (a.abs_type == 2) taper_apply(a, z, y, x, v3, 3);
    own_st(a.vx + i, v3[0]);
    own_st(a.vy + i, v3[1]);
    own_st(a.vz + i, v3[2]);
    // density gradient (src/update_adjv3D.cl:481)
    if (A.grz) {
        own_add(A.grho + i, (double)(-vx * lvx));
        own_add(A.gry + i, (double)(-(vy * lvy)));
        own_add(A.grz + i, (double)(-(vz * lvz)));
    } else {
        const R d = -vx * lvx - vy * lvy - vz * lvz;
        own_add(A.grho + i, (double)d);
    }
}

// ---------------------------------------------------------------------------
// Viscoelastic adjoint steps (L > 0).  The forward state of the step comes
// from the recomputed history (A.st), not from backpropagation, which is
// unstable for the memory-variable recursion.  The adjoint stress update is
// the forward viscoelastic recursion run on the adjoint fields and their own
// memory variables, each l at its own slab (src/update_adjs2D.cl:367-408,
// src/update_adjs3D.cl:549-605 index and write them otherwise); the gradient
// reads the partial increments, as the reference's does.
// ---------------------------------------------------------------------------
// the forward state q of the step at cell i, stored as S
template <typename R, typename S>
__device__ __forceinline__ R hst(const AdjArgs<R> &A, int q, long i) {
    if constexpr (std::is_same<S, __half>::value) return (R)__half2float(((const S *)A.st[q])[i]);
    else return (R)((const S *)A.st[q])[i];
}
template <typename R, typename S>
__device__ __forceinline__ R hst_own(const AdjArgs<R> &A, int q, long i) {
    if constexpr (std::is_same<S, __half>::value)
        return (R)__half2float(__ushort_as_half(own_ld((const unsigned short *)A.st[q] + i)));
    else return (R)own_ld((const S *)A.st[q] + i);
}

template <typename R, typename S, int H, int L>
__global__ __launch_bounds__(256) void k_adjh_s2d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &a = A.r;
    const Layout &g = a.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const R *hc = a.hc;
    R vxx = Dm<H>(a.vx, i, sx, hc), vxz = Dp<H>(a.vx, i, 1, hc);
    R vzz = Dm<H>(a.vz, i, 1, hc), vzx = Dp<H>(a.vz, i, sx, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const long ip = (long)x * 2 * NAB + k;
            pml_half(a, 2, vxz, a.psi_vx_z + ip, bot ? 2 * NAB - 1 - k : k);
            pml_full(a, 2, vzz, a.psi_vz_z + ip, bot ? 2 * NAB - k : k);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const long ip = (long)ii * NZ + z;
            pml_full(a, 0, vxx, a.psi_vx_x + ip, x < NAB ? ii : 2 * NAB - ii);
            pml_half(a, 0, vzx, a.psi_vz_x + ip, x < NAB ? ii : 2 * NAB - 1 - ii);
        }
    }
    const R DT2 = a.DT2;
    const R gg = a.cg[i], f = a.cf[i], d = a.cd[i], e = a.ce[i];
    const R fipkp = a.fipkp[i], dipkp = a.dipkp[i];
    const long nf = g.nfield;
    R rxz[L], rxx[L], rzz[L];
    R sumrxz = 0, sumrxx = 0, sumrzz = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const long ir = l * nf + i;
        rxz[l] = a.rxz[ir]; rxx[l] = a.rxx[ir]; rzz[l] = a.rzz[ir];
        sumrxz += rxz[l]; sumrxx += rxx[l]; sumrzz += rzz[l];
    }
    const R lsxz = (fipkp * (vxz + vzx)) + (DT2 * sumrxz);
    const R lsxx = ((gg * (vxx + vzz)) - (f * vzz)) + (DT2 * sumrxx);
    const R lszz = ((gg * (vxx + vzz)) - (f * vxx)) + (DT2 * sumrzz);
    const R oxz = sumrxz, oxx = sumrxx, ozz = sumrzz;
    sumrxz = sumrxx = sumrzz = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const R b = a.bl[l], c = a.cl[l], et = a.eta[l];
        const long ir = l * nf + i;
        rxz[l] = b * (rxz[l] * c - et * (dipkp * (vxz + vzx)));
        rxx[l] = b * (rxx[l] * c - et * ((e * (vxx + vzz)) - (d * vzz)));
        rzz[l] = b * (rzz[l] * c - et * ((e * (vxx + vzz)) - (d * vxx)));
        a.rxz[ir] = rxz[l]; a.rxx[ir] = rxx[l]; a.rzz[ir] = rzz[l];
        sumrxz += rxz[l]; sumrxx += rxx[l]; sumrzz += rzz[l];
    }
    R s3[3] = {a.sxx[i] + (lsxx + (DT2 * sumrxx)), a.szz[i] + (lszz + (DT2 * sumrzz)),
               a.sxz[i] + (lsxz + (DT2 * sumrxz))};
    if (a.abs_type == 2) taper_apply(a, z, 0, x, s3, 3);
    a.sxx[i] = s3[0];
    a.szz[i] = s3[1];
    a.sxz[i] = s3[2];
    // M and mu gradients (src/update_adjs2D.cl:441-468) from the forward state
    const R sxx = hst<R, S>(A, 2, i), szz = hst<R, S>(A, 3, i), sxz = hst<R, S>(A, 4, i);
    const R lM = A.Ms[i], lmu = A.mus[i];
    grad_mmu_2d(A, i, sxx, szz, sxz, lsxx, lszz, lsxz, lM, lmu);
    if (A.gtp) {
        // tau gradients from the whole memory term of the step
        const R vii[2] = {vxx, vzz};
        const R Rm[2] = {DT2 * oxx + DT2 * sumrxx, DT2 * ozz + DT2 * sumrzz};
        R Xp, Xs[2];
        tau_parts_normal<R>(vxx + vzz, vii, Rm, 2, gg, f, d, e, (R)L, a.DT, A.s0, Xp, Xs);
        const R Xxz = tau_part_shear<R>(vxz + vzx, DT2 * oxz + DT2 * sumrxz, fipkp, dipkp,
                                        lsxz + (DT2 * sumrxz), (R)L, a.DT, A.s0);
        grad_tau_2d(A, i, sxx, szz, sxz, lM, lmu, Xp, Xs[0], Xs[1], Xxz);
    }
}

template <typename R, typename S, int H>
__global__ __launch_bounds__(256) void k_adjh_v2d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &a = A.r;
    const Layout &g = a.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, 0, x), sx = g.sx;
    const R *hc = a.hc;
    const R rip = a.rip[i], rkp = a.rkp[i];
    R sxx_x = Dp<H>(a.sxx, i, sx, hc), szz_z = Dp<H>(a.szz, i, 1, hc);
    R sxz_z = Dm<H>(a.sxz, i, 1, hc), sxz_x = Dm<H>(a.sxz, i, sx, hc);
    const int NAB = a.nab, NZ = g.NZ, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const long ip = (long)x * 2 * NAB + k;
            pml_full(a, 2, sxz_z, a.psi_sxz_z + ip, bot ? 2 * NAB - k : k);
            pml_half(a, 2, szz_z, a.psi_szz_z + ip, bot ? 2 * NAB - 1 - k : k);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const long ip = (long)ii * NZ + z;
            pml_half(a, 0, sxx_x, a.psi_sxx_x + ip, x < NAB ? ii : 2 * NAB - 1 - ii);
            pml_full(a, 0, sxz_x, a.psi_sxz_x + ip, x < NAB ? ii : 2 * NAB - ii);
        }
    }
    const R lvx = ((sxx_x + sxz_z) * rip);
    const R lvz = ((szz_z + sxz_x) * rkp);
    R v2[2] = {a.vx[i] + lvx, a.vz[i] + lvz};
    if (a.abs_type == 2) taper_apply(a, z, 0, x, v2, 2);
    a.vx[i] = v2[0];
    a.vz[i] = v2[1];
    // density gradient (src/update_adjv2D.cl:326) from the forward state
    if (A.grz) {
        A.grho[i] += (double)(-hst<R, S>(A, 0, i) * lvx) * A.gscale_v;
        A.grz[i] += (double)(-(hst<R, S>(A, 1, i) * lvz)) * A.gscale_v;
    } else {
        const R d = -hst<R, S>(A, 0, i) * lvx - hst<R, S>(A, 1, i) * lvz;
        A.grho[i] += (double)d * A.gscale_v;
    }
}

template <typename R, typename S, int H, int L>
__global__ __launch_bounds__(256) void k_adjh_s3d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &a = A.r;
    const Layout &g = a.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, y, x), sx = g.sx, sy = g.sy;
    const R *hc = a.hc;
    R vxx = Dm<H>(a.vx, i, sx, hc), vxy = Dp<H>(a.vx, i, sy, hc), vxz = Dp<H>(a.vx, i, 1, hc);
    R vyx = Dp<H>(a.vy, i, sx, hc), vyy = Dm<H>(a.vy, i, sy, hc), vyz = Dp<H>(a.vy, i, 1, hc);
    R vzx = Dp<H>(a.vz, i, sx, hc), vzy = Dp<H>(a.vz, i, sy, hc), vzz = Dm<H>(a.vz, i, 1, hc);
    const int NAB = a.nab, NZ = g.NZ, NY = g.NY, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int ki = bot ? 2 * NAB - k : k, kh = bot ? 2 * NAB - 1 - k : k;
            const long ip = (long)x * NY * 2 * NAB + (long)y * 2 * NAB + k;
            pml_half(a, 2, vxz, a.psi_vx_z + ip, kh);
            pml_half(a, 2, vyz, a.psi_vy_z + ip, kh);
            pml_full(a, 2, vzz, a.psi_vz_z + ip, ki);
        }
        if (y < NAB || y > NY - NAB - 1) {
            const int j = y < NAB ? y : y - NY + 2 * NAB;
            const int ji = y < NAB ? j : 2 * NAB - j, jh = y < NAB ? j : 2 * NAB - 1 - j;
            const long ip = (long)x * 2 * NAB * NZ + (long)j * NZ + z;
            pml_half(a, 1, vxy, a.psi_vx_y + ip, jh);
            pml_full(a, 1, vyy, a.psi_vy_y + ip, ji);
            pml_half(a, 1, vzy, a.psi_vz_y + ip, jh);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            const long ip = (long)ii * NY * NZ + (long)y * NZ + z;
            pml_full(a, 0, vxx, a.psi_vx_x + ip, iI);
            pml_half(a, 0, vyx, a.psi_vy_x + ip, iH);
            pml_half(a, 0, vzx, a.psi_vz_x + ip, iH);
        }
    }
    const R vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
    const R vxxyyzz = vxx + vyy + vzz, vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
    const R DT2 = a.DT2;
    const R fipjp = own_ld(a.fipjp + i), fjpkp = own_ld(a.fjpkp + i), fipkp = own_ld(a.fipkp + i);
    const R gg = own_ld(a.cg + i), f = own_ld(a.cf + i), d = own_ld(a.cd + i), e = own_ld(a.ce + i);
    const R dipjp = own_ld(a.dipjp + i), djpkp = own_ld(a.djpkp + i), dipkp = own_ld(a.dipkp + i);
    const long nf = g.nfield;
    R rxy[L], ryz[L], rxz[L], rxx[L], ryy[L], rzz[L];
    R sumrxy = 0, sumryz = 0, sumrxz = 0, sumrxx = 0, sumryy = 0, sumrzz = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const long ir = l * nf + i;
        rxy[l] = own_ld(a.rxy + ir); ryz[l] = own_ld(a.ryz + ir); rxz[l] = own_ld(a.rxz + ir);
        rxx[l] = own_ld(a.rxx + ir); ryy[l] = own_ld(a.ryy + ir); rzz[l] = own_ld(a.rzz + ir);
        sumrxy += rxy[l]; sumryz += ryz[l]; sumrxz += rxz[l];
        sumrxx += rxx[l]; sumryy += ryy[l]; sumrzz += rzz[l];
    }
    const R lsxy = (fipjp * vxyyx) + (DT2 * sumrxy);
    const R lsyz = (fjpkp * vyzzy) + (DT2 * sumryz);
    const R lsxz = (fipkp * vxzzx) + (DT2 * sumrxz);
    const R lsxx = ((gg * vxxyyzz) - (f * vyyzz)) + (DT2 * sumrxx);
    const R lsyy = ((gg * vxxyyzz) - (f * vxxzz)) + (DT2 * sumryy);
    const R lszz = ((gg * vxxyyzz) - (f * vxxyy)) + (DT2 * sumrzz);
    const R oxy = sumrxy, oyz = sumryz, oxz = sumrxz, oxx = sumrxx, oyy = sumryy, ozz = sumrzz;
    sumrxy = sumryz = sumrxz = sumrxx = sumryy = sumrzz = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const R b = a.bl[l], c = a.cl[l], et = a.eta[l];
        const long ir = l * nf + i;
        rxy[l] = b * (rxy[l] * c - et * (dipjp * vxyyx));
        ryz[l] = b * (ryz[l] * c - et * (djpkp * vyzzy));
        rxz[l] = b * (rxz[l] * c - et * (dipkp * vxzzx));
        rxx[l] = b * (rxx[l] * c - et * ((e * vxxyyzz) - (d * vyyzz)));
        ryy[l] = b * (ryy[l] * c - et * ((e * vxxyyzz) - (d * vxxzz)));
        rzz[l] = b * (rzz[l] * c - et * ((e * vxxyyzz) - (d * vxxyy)));
        own_st(a.rxy + ir, rxy[l]); own_st(a.ryz + ir, ryz[l]); own_st(a.rxz + ir, rxz[l]);
        own_st(a.rxx + ir, rxx[l]); own_st(a.ryy + ir, ryy[l]); own_st(a.rzz + ir, rzz[l]);
        sumrxy += rxy[l]; sumryz += ryz[l]; sumrxz += rxz[l];
        sumrxx += rxx[l]; sumryy += ryy[l]; sumrzz += rzz[l];
    }
    R s6[6] = {own_ld(a.sxy + i) + (lsxy + (DT2 * sumrxy)), own_ld(a.syz + i) + (lsyz + (DT2 * sumryz)),
               own_ld(a.sxz + i) + (lsxz + (DT2 * sumrxz)), own_ld(a.sxx + i) + (lsxx + (DT2 * sumrxx)),
               own_ld(a.syy + i) + (lsyy + (DT2 * sumryy)), own_ld(a.szz + i) + (lszz + (DT2 * sumrzz))};
    if (a.abs_type == 2) taper_apply(a, z, y, x, s6, 6);
    own_st(a.sxy + i, s6[0]);
    own_st(a.syz + i, s6[1]);
    own_st(a.sxz + i, s6[2]);
    own_st(a.sxx + i, s6[3]);
    own_st(a.syy + i, s6[4]);
    own_st(a.szz + i, s6[5]);
    // M and mu gradients (src/update_adjs3D.cl:678-710) from the forward state
    const R sxx = hst_own<R, S>(A, 3, i), syy = hst_own<R, S>(A, 4, i), szz = hst_own<R, S>(A, 5, i);
    const R sxy = hst_own<R, S>(A, 6, i), syz = hst_own<R, S>(A, 7, i), sxz = hst_own<R, S>(A, 8, i);
    const R lM = own_ld(A.Ms + i), lmu = own_ld(A.mus + i);
    grad_mmu_3d(A, i, sxx, syy, szz, sxy, syz, sxz, lsxx, lsyy, lszz, lsxy, lsyz, lsxz, lM, lmu);
    if (A.gtp) {
        // tau gradients from the whole memory term of the step
        const R vii[3] = {vxx, vyy, vzz};
        const R Rm[3] = {DT2 * oxx + DT2 * sumrxx, DT2 * oyy + DT2 * sumryy,
                         DT2 * ozz + DT2 * sumrzz};
        R Xp, X[6];
        const R Lr = (R)L;
        tau_parts_normal<R>(vxxyyzz, vii, Rm, 3, gg, f, d, e, Lr, a.DT, A.s0, Xp, X);
        X[3] = tau_part_shear<R>(vxyyx, DT2 * oxy + DT2 * sumrxy, fipjp, dipjp,
                                 lsxy + (DT2 * sumrxy), Lr, a.DT, A.s0);
        X[4] = tau_part_shear<R>(vyzzy, DT2 * oyz + DT2 * sumryz, fjpkp, djpkp,
                                 lsyz + (DT2 * sumryz), Lr, a.DT, A.s0);
        X[5] = tau_part_shear<R>(vxzzx, DT2 * oxz + DT2 * sumrxz, fipkp, dipkp,
                                 lsxz + (DT2 * sumrxz), Lr, a.DT, A.s0);
        grad_tau_3d(A, i, sxx, syy, szz, sxy, syz, sxz, lM, lmu, Xp, X);
    }
}

template <typename R, typename S, int H>
__global__ __launch_bounds__(256) void k_adjh_v3d(AdjArgs<R> A, int nbz, unsigned nb) {
    const FDArgs<R> &a = A.r;
    const Layout &g = a.g;
    int z, y, x;
    if (!adj_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, y, x), sx = g.sx, sy = g.sy;
    const R *hc = a.hc;
    const R rip = own_ld(a.rip + i), rjp = own_ld(a.rjp + i), rkp = own_ld(a.rkp + i);
    R szz_z = Dp<H>(a.szz, i, 1, hc), sxx_x = Dp<H>(a.sxx, i, sx, hc);
    R sxz_x = Dm<H>(a.sxz, i, sx, hc), sxz_z = Dm<H>(a.sxz, i, 1, hc);
    R syz_y = Dm<H>(a.syz, i, sy, hc), syz_z = Dm<H>(a.syz, i, 1, hc);
    R syy_y = Dp<H>(a.syy, i, sy, hc), sxy_x = Dm<H>(a.sxy, i, sx, hc);
    R sxy_y = Dm<H>(a.sxy, i, sy, hc);
    const int NAB = a.nab, NZ = g.NZ, NY = g.NY, NX = g.NX;
    if (a.abs_type == 1) {
        if (z > NZ - NAB - 1 || (!a.freesurf && z < NAB)) {
            const bool bot = z > NZ - NAB - 1;
            const int k = bot ? z - NZ + 2 * NAB : z;
            const int ki = bot ? 2 * NAB - k : k, kh = bot ? 2 * NAB - 1 - k : k;
            const long ip = (long)x * NY * 2 * NAB + (long)y * 2 * NAB + k;
            pml_full(a, 2, sxz_z, a.psi_sxz_z + ip, ki);
            pml_full(a, 2, syz_z, a.psi_syz_z + ip, ki);
            pml_half(a, 2, szz_z, a.psi_szz_z + ip, kh);
        }
        if (y < NAB || y > NY - NAB - 1) {
            const int j = y < NAB ? y : y - NY + 2 * NAB;
            const int ji = y < NAB ? j : 2 * NAB - j, jh = y < NAB ? j : 2 * NAB - 1 - j;
            const long ip = (long)x * 2 * NAB * NZ + (long)j * NZ + z;
            pml_full(a, 1, sxy_y, a.psi_sxy_y + ip, ji);
            pml_half(a, 1, syy_y, a.psi_syy_y + ip, jh);
            pml_full(a, 1, syz_y, a.psi_syz_y + ip, ji);
        }
        if (x < NAB || x > NX - NAB - 1) {
            const int ii = x < NAB ? x : x - NX + 2 * NAB;
            const int iI = x < NAB ? ii : 2 * NAB - ii, iH = x < NAB ? ii : 2 * NAB - 1 - ii;
            const long ip = (long)ii * NY * NZ + (long)y * NZ + z;
            pml_half(a, 0, sxx_x, a.psi_sxx_x + ip, iH);
            pml_full(a, 0, sxy_x, a.psi_sxy_x + ip, iI);
            pml_full(a, 0, sxz_x, a.psi_sxz_x + ip, iI);
        }
    }
    const R lvx = ((sxx_x + sxy_y + sxz_z) * rip);
    const R lvy = ((syy_y + sxy_x + syz_z) * rjp);
    const R lvz = ((szz_z + sxz_x + syz_y) * rkp);
    R v3[3] = {own_ld(a.vx + i) + lvx, own_ld(a.vy + i) + lvy, own_ld(a.vz + i) + lvz};
    if (a.abs_type == 2) taper_apply(a, z, y, x, v3, 3);
    own_st(a.vx + i, v3[0]);
    own_st(a.vy + i, v3[1]);
    own_st(a.vz + i, v3[2]);
    // density gradient (src/update_adjv3D.cl:481) from the forward state
    if (A.grz) {
        own_add(A.grho + i, (double)(-hst_own<R, S>(A, 0, i) * lvx) * A.gscale_v);
        own_add(A.gry + i, (double)(-(hst_own<R, S>(A, 1, i) * lvy)) * A.gscale_v);
        own_add(A.grz + i, (double)(-(hst_own<R, S>(A, 2, i) * lvz)) * A.gscale_v);
    } else {
        const R d = -hst_own<R, S>(A, 0, i) * lvx - hst_own<R, S>(A, 1, i) * lvy -
                    hst_own<R, S>(A, 2, i) * lvz;
        own_add(A.grho + i, (double)d * A.gscale_v);
    }
}

// ---------------------------------------------------------------------------
// Boundary saving and adjoint sources
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_savebnd(FDArgs<R> f, Shell sh, R *bnd, const int *tb, long per) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= sh.n) return;
    if (tb) bnd += (long)*tb * per;   // replayed step graphs: a device step counter
    int k, j, i;
    sh.cell(gid, k, j, i);
    const int H = f.g.H;
    const long c = f.g.idx(k - H, f.g.nd == 3 ? j - H : 0, i - H);
    const long n = sh.n;
    if (f.g.nd == 3) {
        const R *v[9] = {f.vx, f.vy, f.vz, f.sxx, f.syy, f.szz, f.sxy, f.syz, f.sxz};
#pragma unroll
        for (int q = 0; q < 9; ++q) bnd[q * n + gid] = v[q][c];
    } else {
        const R *v[5] = {f.vx, f.vz, f.sxx, f.szz, f.sxz};
#pragma unroll
        for (int q = 0; q < 5; ++q) bnd[q * n + gid] = v[q][c];
    }
}

template <typename R>
__device__ __forceinline__ void add_residual(const FDArgs<R> &a, int nd_case, int code, int cell,
                                             float v) {
    switch (code) {
        case O_VX: a.vx[cell] += v; break;
        case O_VY: a.vy[cell] += v; break;
        case O_VZ: a.vz[cell] += v; break;
        case O_SXX: a.sxx[cell] += v; break;
        case O_SYY: a.syy[cell] += v; break;
        case O_SZZ: a.szz[cell] += v; break;
        case O_SXZ: a.sxz[cell] += v; break;
        case O_SXY: a.sxy[cell] += v; break;
        case O_SYZ: a.syz[cell] += v; break;
        default: {
            // p spreads over the normal stresses ("out/%d", an int literal)
            const R q = v / (nd_case == 3 ? 3 : 2);
            a.sxx[cell] += q;
            a.szz[cell] += q;
            if (nd_case == 3) a.syy[cell] += q;
        }
    }
}

struct CodeList {
    int n;
    int code[10];
};

// serial: one lane walks every receiver in order (receivers may share cells)
template <typename R>
__global__ void k_residuals(FDArgs<R> a, int nd_case, ShotDev s, CodeList o, const float *res,
                            int NT, int nt, int serial, const int *tbase) {
    if (tbase) nt += *tbase;
    const int r0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int r1 = serial ? s.nrec : r0 + 1;
    if (serial && r0 != 0) return;
    for (int r = r0; r < r1 && r < s.nrec; ++r) {
        const int cell = s.rec_cell[r];
        if (cell < 0) continue;
        for (int q = 0; q < o.n; ++q)
            add_residual(a, nd_case, o.code[q], cell, res[((long)q * s.nrec + r) * NT + nt]);
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the image-method free surface (src/surface2D_adj.cl:35-148,
// src/surface3D_adj.cl:29-180), elastic.  The reference scatters from each
// surface column into its neighbours' velocities; here a first pass
// transforms every column's surface stresses (keeping what the scatter
// needs) and a second gathers, for each column, its neighbours'
// contributions in increasing neighbour order -- the order of a
// column-by-column run, and free of write races.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_surf_adj_a(FDArgs<R> a, R *tmp, int nd, int syy_fix) {
    const Layout &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = nd == 3 ? blockIdx.y : 0;
    if (x >= g.NX) return;
    const long i = g.idx(0, y, x);
    const long col = (long)y * g.NX + x, nc = (long)g.NX * g.NY;
    const R lmu = a.mu[i], lM = a.M[i];
    if (nd == 2) {
        R lsxx = a.sxx[i], lszz = a.szz[i];
        const R temp1 = lsxx, temp2 = lszz;
        const R aa = (R)1.0 / ((R)4.0 * lM * lmu - (R)4.0 * lmu * lmu);
        lsxx = -aa * (lM * temp1 + (-lM + (R)2.0 * lmu) * temp2);
        lszz = -aa * (lM * temp2 + (-lM + (R)2.0 * lmu) * temp1);
        lszz = 0.0;
        tmp[col] = -((lM - lmu * (R)2.0) * (lM - lmu * (R)2.0) / lM);   // hx
        tmp[nc + col] = -(lM - lmu * (R)2.0);                          // hz
        tmp[2 * nc + col] = lsxx;
        a.szz[i] = -(lM * lszz + (lM - (R)2.0 * lmu) * lsxx);
        a.sxx[i] = -(lM * lsxx + (lM - (R)2.0 * lmu) * lszz);
        return;
    }
    R lszz = -a.szz[i], lsxx = -a.sxx[i], lsyy = syy_fix ? -a.syy[i] : -a.sxx[i];
    const R temp1 = lsxx, temp2 = lsyy, temp3 = lszz;
    const R aa = (R)1.0 / ((R)6.0 * lM * lmu - (R)8.0 * lmu * lmu);
    lsxx = aa * ((R)2.0 * (lM - lmu) * temp1 + (-lM + (R)2.0 * lmu) * (temp2 + temp3));
    lsyy = aa * ((R)2.0 * (lM - lmu) * temp2 + (-lM + (R)2.0 * lmu) * (temp1 + temp3));
    lszz = aa * ((R)2.0 * (lM - lmu) * temp3 + (-lM + (R)2.0 * lmu) * (temp1 + temp2));
    lszz = 0.0;
    tmp[col] = -((lM - lmu * (R)2.0) * (lM - lmu * (R)2.0) / lM);       // h
    tmp[nc + col] = -(lM - lmu * (R)2.0);                              // hz
    tmp[2 * nc + col] = lsxx + lsyy;
    a.sxx[i] = -(lM * lsxx + (lM - (R)2.0 * lmu) * (lsyy + lszz));
    a.syy[i] = -(lM * lsyy + (lM - (R)2.0 * lmu) * (lsxx + lszz));
    a.szz[i] = -(lM * lszz + (lM - (R)2.0 * lmu) * (lsxx + lsyy));
}

template <typename R, int H>
__global__ void k_surf_adj_b(FDArgs<R> a, const R *tmp, int nd) {
    const Layout &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = nd == 3 ? blockIdx.y : 0;
    if (x >= g.NX) return;
    const R *hc = a.hc + 1;                 // {HC1..HC6}
    const long nc = (long)g.NX * g.NY;
    const R *th = tmp, *thz = tmp + nc, *ts = tmp + 2 * nc;
    const long i = g.idx(0, y, x);
    // x neighbours (x' = x-H+1 .. x+H), then y neighbours in 3D
    {
        R v = a.vx[i];
        const R rip = a.rip[i];
#pragma unroll
        for (int d = -(H - 1); d <= H; ++d) {
            const int xs = x + d;
            if (xs < 0 || xs >= g.NX) continue;
            const long c = (long)y * g.NX + xs;
            if (d <= 0) v += (hc[-d] * th[c] * ts[c] * rip);
            else v += (-hc[d - 1] * th[c] * ts[c] * rip);
        }
        a.vx[i] = v;
    }
    if (nd == 3) {
        R v = a.vy[i];
        const R rjp = a.rjp[i];
#pragma unroll
        for (int d = -(H - 1); d <= H; ++d) {
            const int ys = y + d;
            if (ys < 0 || ys >= g.NY) continue;
            const long c = (long)ys * g.NX + x;
            if (d <= 0) v += (hc[-d] * th[c] * ts[c] * rjp);
            else v += (-hc[d - 1] * th[c] * ts[c] * rjp);
        }
        a.vy[i] = v;
    }
    const long c = (long)y * g.NX + x;
#pragma unroll
    for (int m = 0; m < H; ++m) a.vz[i + m] += (hc[m] * thz[c] * ts[c] * a.rkp[i + m]);
}

template <typename R>
void grid_adj(const Layout &g, int &nbz, unsigned &nb) {
    nbz = (g.NZ + BZ - 1) / BZ;
    if (g.nd == 3)
        nb = (unsigned)nbz * ((g.NY + ABY - 1) / ABY) * g.NX;
    else
        nb = (unsigned)nbz * ((g.NX + ABY - 1) / ABY);
}

template <typename R, int H>
void run_adj_s(const AdjArgs<R> &a, int nd_case, hipStream_t st) {
    int nbz;
    unsigned nb;
    grid_adj<R>(a.f.g, nbz, nb);
    const dim3 grid(((nb + 7) / 8) * 8), blk(BZ, ABY);
    if (nd_case == 3) hipLaunchKernelGGL((k_adj_s3d<R, H>), grid, blk, 0, st, a, nbz, nb);
    else hipLaunchKernelGGL((k_adj_s2d<R, H>), grid, blk, 0, st, a, nbz, nb);
}
template <typename R, int H>
void run_adj_v(const AdjArgs<R> &a, int nd_case, hipStream_t st) {
    int nbz;
    unsigned nb;
    grid_adj<R>(a.f.g, nbz, nb);
    const dim3 grid(((nb + 7) / 8) * 8), blk(BZ, ABY);
    if (nd_case == 3) hipLaunchKernelGGL((k_adj_v3d<R, H>), grid, blk, 0, st, a, nbz, nb);
    else hipLaunchKernelGGL((k_adj_v2d<R, H>), grid, blk, 0, st, a, nbz, nb);
}

template <typename R, int H>
void run_surf_adj(const FDArgs<R> &r, R *tmp, int nd, int syy_fix, hipStream_t st) {
    const dim3 grid((r.g.NX + 63) / 64, nd == 3 ? r.g.NY : 1);
    hipLaunchKernelGGL((k_surf_adj_a<R>), grid, dim3(64), 0, st, r, tmp, nd, syy_fix);
    hipLaunchKernelGGL((k_surf_adj_b<R, H>), grid, dim3(64), 0, st, r, (const R *)tmp, nd);
}

}  // namespace

template <typename R>
void launch_surface_adj(const FDArgs<R> &r, R *tmp, int nd_case, int syy_fix, void *stream) {
    const int nd = nd_case == 3 ? 3 : 2;
    SEIS_SWITCH_H(r.g.H, (run_surf_adj<R, HH>(r, tmp, nd, syy_fix, (hipStream_t)stream)));
}

template <typename R, typename S, int H, int L>
void run_adjh_s(const AdjArgs<R> &a, int nd_case, hipStream_t st) {
    int nbz;
    unsigned nb;
    grid_adj<R>(a.f.g, nbz, nb);
    const dim3 grid(((nb + 7) / 8) * 8), blk(BZ, ABY);
    if (nd_case == 3) hipLaunchKernelGGL((k_adjh_s3d<R, S, H, L>), grid, blk, 0, st, a, nbz, nb);
    else hipLaunchKernelGGL((k_adjh_s2d<R, S, H, L>), grid, blk, 0, st, a, nbz, nb);
}
template <typename R, typename S, int H>
void run_adjh_v(const AdjArgs<R> &a, int nd_case, hipStream_t st) {
    int nbz;
    unsigned nb;
    grid_adj<R>(a.f.g, nbz, nb);
    const dim3 grid(((nb + 7) / 8) * 8), blk(BZ, ABY);
    if (nd_case == 3) hipLaunchKernelGGL((k_adjh_v3d<R, S, H>), grid, blk, 0, st, a, nbz, nb);
    else hipLaunchKernelGGL((k_adjh_v2d<R, S, H>), grid, blk, 0, st, a, nbz, nb);
}
template <typename R, typename S, int H>
void run_adjh_s_L(const AdjArgs<R> &a, int nd_case, int L, hipStream_t st) {
    switch (L) {
        case 1: run_adjh_s<R, S, H, 1>(a, nd_case, st); break;
        case 2: run_adjh_s<R, S, H, 2>(a, nd_case, st); break;
        case 3: run_adjh_s<R, S, H, 3>(a, nd_case, st); break;
        case 4: run_adjh_s<R, S, H, 4>(a, nd_case, st); break;
        default: break;
    }
}

template <typename R>
void launch_adj_s(const AdjArgs<R> &a, int nd_case, void *stream) {
    SEIS_SWITCH_H(a.f.g.H, (run_adj_s<R, HH>(a, nd_case, (hipStream_t)stream)));
}
template <typename R>
void launch_adj_v(const AdjArgs<R> &a, int nd_case, void *stream) {
    SEIS_SWITCH_H(a.f.g.H, (run_adj_v<R, HH>(a, nd_case, (hipStream_t)stream)));
}
// FP64 runs (R = double) store their history in double: half is ignored
template <typename R>
void launch_adjh_s(const AdjArgs<R> &a, int nd_case, int L, bool half, void *stream) {
    if constexpr (sizeof(R) == 4) {
        if (half) {
            SEIS_SWITCH_H(a.f.g.H,
                          (run_adjh_s_L<R, __half, HH>(a, nd_case, L, (hipStream_t)stream)));
            return;
        }
    }
    SEIS_SWITCH_H(a.f.g.H, (run_adjh_s_L<R, R, HH>(a, nd_case, L, (hipStream_t)stream)));
}
template <typename R>
void launch_adjh_v(const AdjArgs<R> &a, int nd_case, bool half, void *stream) {
    if constexpr (sizeof(R) == 4) {
        if (half) {
            SEIS_SWITCH_H(a.f.g.H, (run_adjh_v<R, __half, HH>(a, nd_case, (hipStream_t)stream)));
            return;
        }
    }
    SEIS_SWITCH_H(a.f.g.H, (run_adjh_v<R, R, HH>(a, nd_case, (hipStream_t)stream)));
}
template <typename R>
void launch_savebnd(const FDArgs<R> &f, const Shell &sh, int nd_case, R *bnd, void *stream,
                    const int *tb, long per) {
    (void)nd_case;
    if (sh.n <= 0) return;
    const unsigned blocks = (unsigned)((sh.n + 255) / 256);
    hipLaunchKernelGGL((k_savebnd<R>), dim3(blocks), dim3(256), 0, (hipStream_t)stream, f, sh, bnd,
                       tb, per);
}
template <typename R>
void launch_residuals(const FDArgs<R> &r, int nd_case, const ShotDev &s, const int *codes,
                      int nout, const float *res, int NT, int nt, int serial, void *stream,
                      const int *tbase) {
    if (s.nrec <= 0) return;
    CodeList o;
    o.n = nout;
    for (int q = 0; q < nout && q < 10; ++q) o.code[q] = codes[q];
    const unsigned blocks = serial ? 1u : (unsigned)((s.nrec + 255) / 256);
    hipLaunchKernelGGL((k_residuals<R>), dim3(blocks), dim3(serial ? 64 : 256), 0,
                       (hipStream_t)stream, r, nd_case, s, o, res, NT, nt, serial, tbase);
}

template void launch_adj_s<float>(const AdjArgs<float> &, int, void *);
template void launch_surface_adj<float>(const FDArgs<float> &, float *, int, int, void *);
template void launch_adj_v<float>(const AdjArgs<float> &, int, void *);
template void launch_adjh_s<float>(const AdjArgs<float> &, int, int, bool, void *);
template void launch_adjh_v<float>(const AdjArgs<float> &, int, bool, void *);
template void launch_savebnd<float>(const FDArgs<float> &, const Shell &, int, float *, void *,
                                    const int *, long);
template void launch_residuals<float>(const FDArgs<float> &, int, const ShotDev &, const int *,
                                      int, const float *, int, int, int, void *, const int *);
// FP64 gradient runs
template void launch_adj_s<double>(const AdjArgs<double> &, int, void *);
template void launch_surface_adj<double>(const FDArgs<double> &, double *, int, int, void *);
template void launch_adj_v<double>(const AdjArgs<double> &, int, void *);
template void launch_adjh_s<double>(const AdjArgs<double> &, int, int, bool, void *);
template void launch_adjh_v<double>(const AdjArgs<double> &, int, bool, void *);
template void launch_savebnd<double>(const FDArgs<double> &, const Shell &, int, double *, void *,
                                     const int *, long);
template void launch_residuals<double>(const FDArgs<double> &, int, const ShotDev &, const int *,
                                       int, const float *, int, int, int, void *, const int *);

}  // namespace seis
