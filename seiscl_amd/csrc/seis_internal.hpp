// Internal types of the MI355X SeisCL engine (not part of the C-ABI).
//
// HBM layout.  Every wavefield, memory variable and kernel coefficient array
// uses one padded layout so that a single linear index serves all arrays of
// a cell:
//   idx(z,y,x) = (x+H)*sx + (y+H)*sy + (z+ZO)      (interior coordinates)
// with H = FDORDER/2 halo cells on each side of y and x, a front pad of ZO
// elements in z chosen so that interior z=0 starts on a 128-byte line, and
// rows of ldz elements rounded up to whole 128-byte lines.  2D uses y = 0,
// sy = 0.  CPML memory arrays keep the compact strip layout of the
// reference (src/update_v3D.cl:23-31).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/seiscl_amd.h"

#if defined(__HIPCC__)
#define SEIS_HD __host__ __device__
#else
#define SEIS_HD
#endif

namespace seis {

constexpr int kMaxL = SEIS_MAXL;
constexpr int kMaxNab = 240;
constexpr int kMaxDftFreqs = 512;   // gradient frequencies of the DFT gradient
constexpr double kPiDft = 3.141592653589793238462643383279502884197169;

struct Layout {
    int nd = 2;            // 2 or 3 (acoustic is 2)
    int NZ = 0, NY = 1, NX = 0;
    int H = 0;             // FDOH
    int ZO = 0;            // z front pad (elements)
    long ldz = 0;          // row length (elements)
    long sy = 0, sx = 0;   // strides of y and x
    long nfield = 0;       // elements per field
    int elem = 4;          // bytes per element of the storage type
    SEIS_HD long idx(int z, int y, int x) const {
        return (long)(x + H) * sx + (long)(y + H) * sy + (z + ZO);
    }
};

Layout make_layout(int nd, int NZ, int NY, int NX, int H, int elem);

// Host-side set-up products (all float, reference semantics).
struct Setup {
    float hc[7] = {0};
    float DT = 0, DT2 = 0, DH = 0;
    float eta[kMaxL] = {0}, bl[kMaxL] = {0}, cl[kMaxL] = {0}, bs[kMaxL] = {0};
    // CPML: [axis x,y,z][K, a, b, K_half, a_half, b_half][2*nab]
    std::vector<float> cpml;
    std::vector<float> taper;
};

int build_setup(const seis_csts &c, Setup &s, std::string &err);

// Reference parameter transforms (src/assign_modeling_case.c:305-535),
// outputs in the reference's unpadded layout.
struct RefParams {
    std::vector<float> M, mu, rho, rip, rjp, rkp, muipkp, muipjp, mujpkp;
    std::vector<float> taup, taus, tausipkp, tausipjp, tausjpkp;
    // viscoelastic cases: the moduli at the reference frequency f0, before
    // the relaxation division (what the transform gives with L=0)
    std::vector<float> Mf0, muf0;
};
int transform_params(const seis_csts &c, const float *p0, const float *p1,
                     const float *rho, const float *taup, const float *taus,
                     RefParams &out, std::string &err);
int check_stability(const seis_csts &c, const RefParams &p, float *dtstable);
// Unpadded reference array (z fastest) -> padded layout, times mul (a power
// of two: exact), converted to T.
template <typename T>
void to_padded(const Layout &g, const float *ref, float mul, std::vector<T> &out);
// Derived viscoelastic coefficients of the stress update with the
// reference's expressions (src/update_s3D.cl:324-340, src/update_s2D.cl:
// 252-259, src/update_s2D_acc.cl:343-346): g, f, d, e (f, d for ND 2/3).
void visco_coeffs(const seis_csts &c, const RefParams &p, float DT, std::vector<float> &g,
                  std::vector<float> &f, std::vector<float> &d, std::vector<float> &e);
// s0 = sum_l w0^2 t_l^2 / (1 + w0^2 t_l^2), t_l = dt/eta_l: the relaxed
// moduli are M_f0 / (1 + s0 tau) (src/assign_modeling_case.c:415-437).
float relax_sum(const seis_csts &c);
// The staggered shear pair: f_ij = mu_ij (1 + L tau_ij), d_ij = mu_ij tau_ij / DT.
void visco_pair(const std::vector<float> &mu, const std::vector<float> &ts, int L, float DT,
                std::vector<float> &ff, std::vector<float> &dd);
// Gradients through the staggered averages of transform_params (gradfix
// FIX_AVG), unpadded layout.  sh: per-node shear products (xz, xy, yz), rh:
// per-node density terms (vx, vy, vz nodes); gmu gets the chained shear terms
// added (with the moduli mus of the gradient formulas), grho is replaced by
// the chained density terms.
void chain_staggered(const seis_csts &c, const RefParams &p, const std::vector<float> &mus,
                     const std::vector<double> *sh, const std::vector<double> *rh, double *gmu,
                     double *grho);

// Output names of seisout (SeisCL.py to_load_names order).
std::vector<std::string> out_names(const seis_csts &c);

// ---------------------------------------------------------------------------
// Kernel argument block (passed by value).  R = compute/storage real type.
// ---------------------------------------------------------------------------
template <typename R>
struct FDArgs {
    Layout g;
    R hc[7];
    R DT, DT2;
    int L;
    R eta[kMaxL], bl[kMaxL], cl[kMaxL];
    int abs_type, freesurf, nab;
    int fuse_surface;                    // apply the free surface inside update_s
    R bs[kMaxL];                         // eta/(1+eta/2), surface memory update
    // wavefields
    R *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *syz, *sxz, *p;
    // memory variables: L consecutive fields each
    R *rxx, *ryy, *rzz, *rxy, *ryz, *rxz, *rp;
    // kernel coefficients (padded layout)
    const R *rip, *rjp, *rkp;
    const R *M, *mu, *muipkp, *muipjp, *mujpkp;        // elastic
    const R *cg, *cf, *cd, *ce;                          // visco: g, f, d, e
    const R *fipkp, *fipjp, *fjpkp, *dipkp, *dipjp, *djpkp;
    // CPML coefficients [3][6][2*nab] and Cerjan taper [nab]
    const R *cpml;
    const R *taper;
    // CPML memory (compact strips)
    R *psi_sxx_x, *psi_sxy_x, *psi_sxz_x, *psi_sxy_y, *psi_syy_y, *psi_syz_y,
        *psi_sxz_z, *psi_syz_z, *psi_szz_z;
    R *psi_vx_x, *psi_vy_x, *psi_vz_x, *psi_vx_y, *psi_vy_y, *psi_vz_y,
        *psi_vx_z, *psi_vy_z, *psi_vz_z;
    R *psi_p_x, *psi_p_z;
};

// Shot description on the device.
struct ShotDev {
    int nsrc = 0, nrec = 0;
    int *src_cell = nullptr;   // [nsrc] linear cell index, -1 = skipped
    int *src_type = nullptr;   // [nsrc]
    int *src_k = nullptr;      // [nsrc] padded z index (free-surface rule)
    float *src = nullptr;      // [nsrc][NT]
    int *rec_cell = nullptr;   // [nrec]  -1 = outside
    float *out = nullptr;      // [NT][nout][nrec] (time-major on the device)
};

// Launchers implemented in the .hip kernel files.  S: storage type of the
// wavefields and memory variables (R, or __half with R = float).  With FP16
// storage, 3D runs fuse the free surface into the stress update and
// launch_surface does nothing for them.
template <typename R, typename S = R>
void launch_update_v(const FDArgs<R> &a, int nd_case, void *stream);
template <typename R, typename S = R>
void launch_update_s(const FDArgs<R> &a, int nd_case, void *stream);
template <typename R, typename S = R>
void launch_surface(const FDArgs<R> &a, int nd_case, int pdir, void *stream);
// S: storage type of wavefields and memory variables (float, __half or
// double); R: compute and coefficient type.
// tbase (may be NULL): device step counter added to nt (replayed step graphs)
template <typename R, typename S>
void launch_sources(const FDArgs<R> &a, int nd_case, const ShotDev &s, int NT,
                    int nt, int pdir, void *stream, const int *tbase = nullptr);
template <typename R, typename S>
void launch_varsout(const FDArgs<R> &a, int nd_case, const ShotDev &s,
                    const int *out_codes, const float *unscale, int nout, int nt,
                    void *stream, const int *tbase = nullptr);
// varsout of step t_rec then the sources of step t_src in one launch (steps
// outside [0, NT) skipped), each plus *tbase when tbase is given
template <typename R, typename S>
void launch_io(const FDArgs<R> &a, int nd_case, const ShotDev &s, const int *out_codes,
               const float *unscale, int nout, int NT, int t_rec, int t_src, void *stream,
               const int *tbase = nullptr);
// *t = v (add 0) or *t += v (add 1), one thread on the stream
void launch_tcounter(int *t, int v, int add, void *stream);
void launch_fill(void *ptr, size_t bytes, void *stream);
template <typename R, typename S>
void launch_v3d(const FDArgs<R> &a, void *stream);
template <typename R, typename S>
void launch_s3d(const FDArgs<R> &a, void *stream);

// Output codes used by the receiver kernel.
enum OutCode { O_VX, O_VY, O_VZ, O_SXX, O_SYY, O_SZZ, O_SXZ, O_SXY, O_SYZ, O_P };

// ---------------------------------------------------------------------------
// Batched serving: one slot per modeling job (2D, L=0, FP32, no free surface).
// ---------------------------------------------------------------------------
struct BatchSlot {
    FDArgs<float> a;
    ShotDev shot;
    int active, nd_case, NT, tmax;
    int start;          // the batch step that runs the job's step 0: step t of
                        // the job is batch step start + t
    int nout, codes[10];
};
// every slot of a batch grouped by kernel class (acoustic or elastic, FD
// order), with the prefix sum of their block counts; a block of a batch pass
// updates kBatchTiles consecutive 64x4 tiles of one slot
#ifndef SEIS_BATCH_TILES
#define SEIS_BATCH_TILES 4
#endif
constexpr int kBatchTiles = SEIS_BATCH_TILES;
struct BatchTable {
    int n;
    unsigned nblocks;
    int tiles = kBatchTiles;   // consecutive tiles per block
    const int *ids;          // [n] slot indices (device)
    const unsigned *block0;  // [n+1] first block of each slot (device)
};
// Batch steps step0 .. step0+nsteps-1: per step one source launch, one
// velocity and one stress launch for all slots, one receiver launch.
void launch_batch_steps(const BatchSlot *dslots, int nslots, const BatchTable &tab, int maxrec,
                        int step0, int nsteps, void *stream);

// ---------------------------------------------------------------------------
// Gradient by backpropagation (BACK_PROP_TYPE=1).
// ---------------------------------------------------------------------------
// Checkpoint plan of a viscoelastic gradient (see plan_checkpoints): L
// levels, spacing K[l] of the level-l checkpoints (K[L-1]: steps of
// recomputed history), ratio n = K[l] / K[l+1], n0 level-0 checkpoints.
constexpr int kMaxLevels = 4;
struct CkptPlan {
    int L = 0, n = 1, n0 = 0;
    int K[kMaxLevels] = {0, 0, 0, 0};
    size_t bytes = 0;
};

// Corrections of reference defects, selected by seis_csts.gradfix.
enum GradFix {
    FIX_RES_SCALE = 1, FIX_ADJ_CPML = 2, FIX_RHO_SCALE = 4, FIX_SURF_SYY = 8, FIX_VISCO_F0 = 16,
    FIX_AVG = 32, FIX_RMSNORM = 64, FIX_HESS = 128
};

// Boundary shell saved during the forward run and re-injected while the
// forward wavefield runs backwards: FDOH cells thick, on the inner face of
// the absorbing layer, enumerated front, back, [left, right,] down, up as
// savebnd does (src/savebnd2D.cl:43-223, src/savebnd3D.cl:38-284); the
// inverse lookup is inject_ind (src/header_injectbnd.cl:25-232, :238-398).
// Coordinates (k, j, i) are the reference's padded ones (interior + FDOH).
struct Shell {
    int nd = 2, F = 0, NAB = 0;
    int NXbnd = 0, NYbnd = 0, NZbnd = 0, lbnd = 0, lbnds = 0;
    long n = 0;

    SEIS_HD void cell(long gid, int &k, int &j, int &i) const {
        const long NXb = NXbnd, NYb = NYbnd, NZb = NZbnd;
        long g;
        j = 0;
        if (nd == 3) {
            const long fb = NYb * NZb * F, lr = NZb * (NXb - 2 * F) * F,
                       du = (NYb - 2 * F) * (NXb - 2 * F) * F;
            if (gid < fb) {
                i = (int)(gid / (NYb * NZb)) + lbnd;
                j = (int)((gid / NZb) % NYb) + lbnd;
                k = (int)(gid % NZb) + lbnds;
            } else if (gid < 2 * fb) {
                g = gid - fb;
                i = (int)(g / (NYb * NZb) + NXb + NAB);
                j = (int)((g / NZb) % NYb) + lbnd;
                k = (int)(g % NZb) + lbnds;
            } else if (gid < 2 * fb + lr) {
                g = gid - 2 * fb;
                i = (int)(g / (NZb * F)) + lbnd + F;
                j = (int)((g / NZb) % F) + lbnd;
                k = (int)(g % NZb) + lbnds;
            } else if (gid < 2 * fb + 2 * lr) {
                g = gid - 2 * fb - lr;
                i = (int)(g / (NZb * F)) + lbnd + F;
                j = (int)((g / NZb) % F + NYb + NAB);
                k = (int)(g % NZb) + lbnds;
            } else if (gid < 2 * fb + 2 * lr + du) {
                g = gid - 2 * fb - 2 * lr;
                i = (int)(g / (F * (NYb - 2 * F))) + lbnd + F;
                j = (int)((g / F) % (NYb - 2 * F)) + lbnd + F;
                k = (int)(g % F + NZb) + lbnds - F;
            } else {
                g = gid - 2 * fb - 2 * lr - du;
                i = (int)(g / (F * (NYb - 2 * F))) + lbnd + F;
                j = (int)((g / F) % (NYb - 2 * F)) + lbnd + F;
                k = (int)(g % F) + lbnds;
            }
            return;
        }
        const long fb = NZb * F, du = (NXb - 2 * F) * F;
        if (gid < fb) {
            i = (int)(gid / NZb) + lbnd;
            k = (int)(gid % NZb) + lbnds;
        } else if (gid < 2 * fb) {
            g = gid - fb;
            i = (int)(g / NZb + NXb + NAB);
            k = (int)(g % NZb) + lbnds;
        } else if (gid < 2 * fb + du) {
            g = gid - 2 * fb;
            i = (int)(g % (NXb - 2 * F)) + lbnd + F;
            k = (int)(g / (NXb - 2 * F) + NZb) + lbnds - F;
        } else {
            g = gid - 2 * fb - du;
            i = (int)(g % (NXb - 2 * F)) + lbnd + F;
            k = (int)(g / (NXb - 2 * F)) + lbnds;
        }
    }

    // shell entry of padded cell (k, j, i), or -1
    SEIS_HD long inject(int k, int j, int i) const {
        const long NXb = NXbnd, NYb = NYbnd, NZb = NZbnd;
        i -= lbnd;
        k -= lbnds;
        if (nd == 3) {
            j -= lbnd;
            if ((k > F - 1 && k < NZb - F) && (j > F - 1 && j < NYb - F) &&
                (i > F - 1 && i < NXb - F))
                return -1;
            if (k < 0 || k > NZb - 1 || j < 0 || j > NYb - 1 || i < 0 || i > NXb - 1) return -1;
            if (i < F) return (long)i * NYb * NZb + (long)j * NZb + k;
            if (i > NXb - 1 - F) return NYb * NZb * F + (long)(i - NXb + F) * NYb * NZb + (long)j * NZb + k;
            if (j < F) return NYb * NZb * F * 2 + (long)(i - F) * F * NZb + (long)j * NZb + k;
            if (j > NYb - 1 - F)
                return NYb * NZb * F * 2 + (NXb - 2 * F) * NZb * F + (long)(i - F) * F * NZb +
                       (long)(j - NYb + F) * NZb + k;
            if (k < F)
                return NYb * NZb * F * 2 + (NXb - 2 * F) * NZb * F * 2 +
                       (NXb - 2 * F) * (NYb - 2 * F) * F + (long)(i - F) * (NYb - 2 * F) * F +
                       (long)(j - F) * F + k;
            return NYb * NZb * F * 2 + (NXb - 2 * F) * NZb * F * 2 +
                   (long)(i - F) * (NYb - 2 * F) * F + (long)(j - F) * F + (k - NZb + F);
        }
        if ((k > F - 1 && k < NZb - F) && (i > F - 1 && i < NXb - F)) return -1;
        if (k < 0 || k > NZb - 1 || i < 0 || i > NXb - 1) return -1;
        if (i < F) return (long)i * NZb + k;
        if (i > NXb - 1 - F) return NZb * F + (long)(i - NXb + F) * NZb + k;
        if (k < F) return NZb * F * 2 + (NXb - 2 * F) * F + (i - F) + (long)k * (NXb - 2 * F);
        return NZb * F * 2 + (i - F) + (long)(k - NZb + F) * (NXb - 2 * F);
    }
};
Shell make_shell(const Layout &g, int nab, int freesurf);

// Arguments of the adjoint kernels: the forward run being backpropagated
// (f), the adjoint run (r: its own wavefields and CPML memory, the same
// model), the gradient accumulators and the boundary of the previous step.
template <typename R>
struct AdjArgs {
    FDArgs<R> f, r;
    double *gM, *gmu, *grho;     // padded layout, one double per cell
    const R *bnd;                // [nvars][n] shell values of step t-1
    Shell sh;
    // viscoelastic runs: the forward state of the step from the recomputed
    // history (2D: vx vz sxx szz sxz; 3D: vx vy vz sxx syy szz sxy syz sxz),
    // in the model's storage type and units, and the moduli of the gradient
    // formulas (padded layout)
    const void *st[9];
    const R *Ms, *mus;
    // factors taking the accumulated M/mu and rho increments back to
    // physical units (1 unless the wavefields are stored scaled)
    double gscale_s = 1, gscale_v = 1;
    // restype 1 (RTM): the M image without the moduli, no mu term
    // (src/update_adjs2D.cl:459-468, src/update_adjs3D.cl:701-710)
    int rtm = 0;
    // approximate Hessian of M and mu (Hout), double, padded layout, or null;
    // hscale takes it to physical units; hfix: mu's uses the Hessian's own
    // M term (the reference's 2D kernel subtracts the gradient's)
    double *hM = nullptr, *hmu = nullptr;
    double hscale = 1;
    int hfix = 1;
    // viscoelastic runs: tau_p and tau_s gradient accumulators (double,
    // padded layout, scaled by gscale_s like M's) or null, and the
    // relaxation sum s0 of the transform (relax_sum)
    double *gtp = nullptr, *gts = nullptr;
    R s0 = 0;
    // gradfix FIX_AVG: the terms of parameters that reach the update through
    // staggered averages, kept per staggered node for the host to chain
    // through the averages (chain_staggered): the shear-stress products
    // sigma_ij times its adjoint increment (xz, xy, yz; not divided by mu^2)
    // and the density terms of vy and vz (vx's stays in grho), scaled like
    // the accumulators they leave.  Null: the reference's attribution of
    // every term to the cell of the same index.
    double *gsh[3] = {nullptr, nullptr, nullptr};
    double *gry = nullptr, *grz = nullptr;
    // replayed step graphs: the shell read is bnd + (*tb + bofs) * bper
    const int *tb = nullptr;
    long bofs = 0, bper = 0;
};

// DFT-domain gradient (dft.hip).  The arrays whose spectra are kept:
// velocities, stresses, then each memory-variable component's L slabs.
struct DftArrays {
    int n = 0;
    const float *p[9 + 6 * kMaxL] = {};
};
// F (complex float pairs) [n][nf][nfield] += fact[f] u; fact: nf pairs
void launch_dft_acc(const DftArrays &q, float *F, const float *fact, int nf, const Layout &g,
                    void *stream);
struct DftGrad {
    const void *Ff = nullptr, *Fa = nullptr;   // forward, adjoint spectra
    const float *par = nullptr;   // [5][N] unpadded: M mu taup taus rho (buoyancy)
    const float *w = nullptr;     // [nf] angular frequencies 2 pi df bin
    double *gM = nullptr, *gmu = nullptr, *grho = nullptr, *gtp = nullptr, *gts = nullptr;
    int nf = 0, L = 0, pt = 0, ntnyq = 1;
    float al = 0;
    float tausig[kMaxL] = {};
};
void launch_dft_grad(const DftGrad &d, const Layout &g, void *stream);

template <typename R>
void launch_adj_s(const AdjArgs<R> &a, int nd_case, void *stream);
template <typename R>
void launch_adj_v(const AdjArgs<R> &a, int nd_case, void *stream);
// viscoelastic adjoint steps reading the forward state from A.st (L = 1..4),
// stored as float, or as half when half is set
template <typename R>
void launch_adjh_s(const AdjArgs<R> &a, int nd_case, int L, bool half, void *stream);
template <typename R>
void launch_adjh_v(const AdjArgs<R> &a, int nd_case, bool half, void *stream);
// adjoint free surface on the adjoint wavefields r; tmp: 3 * NX * NY reals
template <typename R>
void launch_surface_adj(const FDArgs<R> &r, R *tmp, int nd_case, int syy_fix, void *stream);
// bnd[q][g] = field_q(shell cell g), q over the shell variables
template <typename R>
// tb (may be NULL): the shell goes to bnd + *tb * per (replayed step graphs)
void launch_savebnd(const FDArgs<R> &f, const Shell &sh, int nd_case, R *bnd, void *stream,
                    const int *tb = nullptr, long per = 0);
// adjoint sources of step nt into the adjoint wavefields; res [nout][nrec][NT]
template <typename R>
// tbase (may be NULL): device step counter added to nt (replayed step graphs)
void launch_residuals(const FDArgs<R> &r, int nd_case, const ShotDev &s, const int *out_codes,
                      int nout, const float *res, int NT, int nt, int serial, void *stream,
                      const int *tbase = nullptr);

// Host side of the gradient (src/residuals.c, src/calc_grad.c:975-1110).
double residuals(const seis_csts &c, int nout, int nrec, const float *out, const float *din,
                 float *res, double *rmsnorm);
void butterworth(float *data, float fcl, float fch, float dt, int NT, int tmax, int ntrace,
                 int order);
int fft_fast_size(int m);
// rms and rmsnorm at the DFT gradient's frequencies (src/residuals.c:364-403):
// |X(k)|^2 of the traces -out + din and din (band-passed like the
// residuals, zero-padded to fft_fast_size(tmax)) at k = (int)(f nfft dt + 1),
// the DFT bin evaluated directly in double.
double dft_rms(const seis_csts &c, int nout, int nrec, const float *out, const float *din,
               const std::vector<float> &freqs, double *rmsnorm);
void res_scale(const seis_csts &c, const RefParams &p, const std::vector<std::string> &outn,
               int nrec, const float *rec_pos, float *res, int at_receiver);
void res_scale_factors(const seis_csts &c, const RefParams &p, const std::vector<std::string> &outn,
                       int nrec, const float *rec_pos, double *fac, int at_receiver);
void transf_hess(const seis_csts &c, const RefParams &p, const std::vector<float> &Ms,
                 const std::vector<float> &mus, const double *hM, const double *hmu,
                 float *out);
void transf_grad(const seis_csts &c, const RefParams &p, const std::vector<float> &Ms,
                 const std::vector<float> &mus, const double *gM, const double *gmu,
                 const double *grho, float *out, int rho_fix);

}  // namespace seis
