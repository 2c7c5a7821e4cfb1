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
};
int transform_params(const seis_csts &c, const float *p0, const float *p1,
                     const float *rho, const float *taup, const float *taus,
                     RefParams &out, std::string &err);
int check_stability(const seis_csts &c, const RefParams &p, float *dtstable);

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

// Launchers implemented in the .hip kernel files.
template <typename R>
void launch_update_v(const FDArgs<R> &a, int nd_case, void *stream);
template <typename R>
void launch_update_s(const FDArgs<R> &a, int nd_case, void *stream);
template <typename R>
void launch_surface(const FDArgs<R> &a, int nd_case, int pdir, void *stream);
template <typename R>
void launch_sources(const FDArgs<R> &a, int nd_case, const ShotDev &s, int NT,
                    int nt, int pdir, void *stream);
template <typename R>
void launch_varsout(const FDArgs<R> &a, int nd_case, const ShotDev &s,
                    const int *out_codes, int nout, int nt, void *stream);
void launch_fill(void *ptr, size_t bytes, void *stream);
template <typename R>
void launch_v3d(const FDArgs<R> &a, void *stream);
template <typename R>
void launch_s3d(const FDArgs<R> &a, void *stream);

// Output codes used by the receiver kernel.
enum OutCode { O_VX, O_VY, O_VZ, O_SXX, O_SYY, O_SZZ, O_SXZ, O_SXY, O_SYZ, O_P };

}  // namespace seis
