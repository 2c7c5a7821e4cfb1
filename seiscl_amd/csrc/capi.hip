// C-ABI runtime: device context, buffers, model set-up and the per-shot
// time loop.  Replaces the reference's Init_OpenCL / clprogram / clbuf layer
// and the forward half of time_stepping (src/time_stepping.c:553-848).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "seis_internal.hpp"

using namespace seis;

namespace {
thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return 1;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return fail(std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

}  // namespace

// The fixed device area of seis_forward_shot_static.  In host code the name
// stands for a host-side shadow, not for device memory: host code uses only
// the device address that hipGetSymbolAddress returns for the current GPU
// (kept per device in g_scratch_dev, set up once by seis_ctx_create).
static __device__ __align__(256) unsigned char scratch[SEIS_STATIC_SHOT_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of this GPU's static area (the device must be current).
int scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return 0;   // the static path refuses later
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return 0;
}

// Seismograms recorded time-major ([NT][nout][nrec], float or double) into
// the boundary's [nout][nrec][NT] floats, in the device area next to them.
// Adjoint sources given on the device ([nout][nrec][NT]): res_scale factors
// applied to t < tmax as (float)((double)r * f), then the velocity scale of
// scaled-unit runs (qv[o], a float product), as the host path does.
__global__ void k_res_device(const float *in, float *out, const double *fac, const float *qv,
                             int nout, int nrec, int NT, int tmax) {
    const long n = (long)nout * nrec * NT;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
         k += (long)gridDim.x * blockDim.x) {
        const long tr = k / NT;
        const int t = (int)(k % NT), o = (int)(tr / nrec);
        float v = in[k];
        if (fac && t < tmax) v = (float)((double)v * fac[tr]);
        if (qv) v = v * qv[o];
        out[k] = v;
    }
}

template <typename T>
__global__ void k_seis_transpose(const T *in, float *out, int NT, int nout, int nrec) {
    const long n = (long)NT * nout * nrec;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
         k += (long)gridDim.x * blockDim.x) {
        const long r = k % nrec, o = (k / nrec) % nout, t = k / ((long)nrec * nout);
        out[(o * nrec + r) * NT + t] = (float)in[k];
    }
}
}  // namespace

struct seis_ctx {
    int device = 0;
    void *scratch = nullptr;        // device address of this GPU's static shot area
    hipStream_t stream = nullptr;   // compute
    hipStream_t xfer = nullptr;     // model set-up uploads and result copies
    size_t bytes = 0;               // bytes handed out and not released
    std::mutex mu;                  // models may be set up on another host thread
    struct Block {
        size_t cap;
        hipEvent_t freed;           // compute-stream point after the last use
    };
    std::map<void *, size_t> allocs;          // live block -> requested bytes
    std::map<void *, size_t> caps;            // every block -> capacity
    std::multimap<size_t, std::pair<void *, hipEvent_t>> cache;   // released blocks

    // Device memory, zero-filled on fill_stream (the compute stream by
    // default).  Released blocks are kept and reused for requests of up to
    // their size and at least half of it: no hipMalloc / hipFree (which
    // synchronises the device) while jobs come and go.
    int alloc(size_t n, void **p, hipStream_t fill_stream = nullptr) {
        *p = nullptr;
        if (n == 0) return 0;
        hipStream_t fs = fill_stream ? fill_stream : stream;
        std::lock_guard<std::mutex> lock(mu);
        HIP_TRY(hipSetDevice(device));
        auto it = cache.lower_bound(n);
        if (it != cache.end() && it->first <= 2 * n) {
            *p = it->second.first;
            // the block's previous user may still be queued on the compute stream
            HIP_TRY(hipStreamWaitEvent(fs, it->second.second, 0));
            (void)hipEventDestroy(it->second.second);
            cache.erase(it);
        } else {
            HIP_TRY(hipMalloc(p, n));
            caps[*p] = n;
        }
        allocs[*p] = n;
        bytes += n;
        launch_fill(*p, n, fs);
        return 0;
    }
    int release(void *p) {
        if (!p) return 0;
        std::lock_guard<std::mutex> lock(mu);
        auto it = allocs.find(p);
        if (it == allocs.end()) return fail("seis_buf_free: unknown pointer");
        bytes -= it->second;
        allocs.erase(it);
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(e, stream));
        cache.emplace(caps[p], std::make_pair(p, e));
        return 0;
    }
};

// Precision modes (seis_csts.FP16): FP32; FP32 storage with the scaled
// units of FP16 (a test of the scaling, results equal FP32 up to
// subnormals); FP16 storage of wavefields and memory variables with FP32
// arithmetic; FP64 storage and arithmetic.
enum Prec { P_FP32 = 0, P_SCALED = 1, P_FP16 = 2, P_FP64 = 64 };

struct seis_model {
    seis_ctx *ctx = nullptr;
    seis_csts c{};
    int nd_case = 2;
    Layout g;
    Setup s;
    FDArgs<float> a{};
    // precision: storage bytes of wavefields (selem) and of coefficients (celem)
    int prec = P_FP32, selem = 4, celem = 4;
    FDArgs<double> ad{};             // the same buffers, FP64 mode
    // scaled units: stored v = av * v, stored stresses = as * sigma, with
    // av / as = q fixed by the model and av chosen per shot from the source
    float q = 1, av = 1, as = 1;
    std::vector<float> unscale;      // per output: 1/av or 1/as
    std::vector<void *> owned;       // device buffers owned by the model
    size_t bytes = 0;
    std::vector<std::string> outn;
    std::vector<int> out_codes;
    // shot
    ShotDev shot;
    int shot_nsrc_cap = 0, shot_nrec_cap = 0;
    bool shot_bound = false;
    // timing
    bool profiling = false;
    std::vector<hipEvent_t> evpool;
    double ms_v = 0, ms_s = 0, ms_loop = 0;
    // host copy of the transformed parameters (residual scaling, gradients)
    RefParams rp;
    bool setting_up = false;         // buffers zero-filled on the transfer stream
    hipEvent_t done_ev = nullptr;    // batch: compute-stream point after the last step
    bool in_batch = false;           // held by a serving batch slot
    // captured forward steps (seis_run_steps, 2D): graph, its key, step counter
    hipGraph_t sg_graph = nullptr;
    hipGraphExec_t sg_exec = nullptr;
    uint64_t sg_key = 0;
    int *sg_t = nullptr;
    // captured steps of the 2D elastic gradient: the forward run with
    // boundary saving and the adjoint run (same step counter)
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        uint64_t key = 0;
        void reset() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
            e = nullptr;
            g = nullptr;
            key = 0;
        }
    } fs_graph, ad_graph;
    std::vector<int> rcell_host;     // receiver cells of the bound shot
    std::vector<float> recpos_host;  // receiver positions of the bound shot [nrec][8]
    // gradient state, allocated on first use
    struct Grad {
        bool ready = false;
        Shell sh;
        int nvb = 0;
        float *bnd = nullptr;       // [tmax][nvb][n]
        long bnd_steps = 0;
        size_t bnd_elem = 4;        // bytes per saved value (8 in FP64 runs)
        FDArgs<float> r{};          // adjoint wavefields (same model)
        double *gM = nullptr, *gmu = nullptr, *grho = nullptr;   // padded layout
        double *hM = nullptr, *hmu = nullptr;                    // Hessian (Hout)
        double *gtp = nullptr, *gts = nullptr;   // tau_p, tau_s (viscoelastic)
        // gradfix FIX_AVG: per-node shear products (xz, xy, yz) and density
        // terms of vy and vz (AdjArgs::gsh, gry, grz)
        double *gsh[3] = {nullptr, nullptr, nullptr};
        double *gry = nullptr, *grz = nullptr;
        float *res = nullptr;       // [nout][nrec][NT] scaled adjoint sources
        float *stmp = nullptr;      // adjoint free surface: 3 * NX * NY
        size_t res_cap = 0;
        double *res_fac = nullptr;  // [nout][nrec] res_scale factors (device residuals)
        size_t res_fac_cap = 0;
        std::vector<float *> own_psi;
        bool saved = false;         // a forward run with boundaries is bound
        // viscoelastic runs: full-state checkpoints every K steps and the
        // recomputed post-source history of one segment of K steps
        bool visco = false;
        CkptPlan plan;              // checkpoint plan of the last forward run
        float *ckpt = nullptr;      // [n1 + n2 - 1][state]
        float *hist = nullptr;      // [K2][nvb][nfield] of the storage type
        size_t ckpt_bytes = 0, hist_bytes = 0;
        size_t mem_limit = 0;       // budget of ckpt + hist; 0: 90% of free memory
        float *Ms = nullptr, *mus = nullptr;   // moduli of the gradient formulas
    } grad;
    // DFT-domain gradient (seis_grad_dft): sampling of the spectra, the
    // factor table, forward and adjoint spectra, the cell parameters
    struct Dft {
        int nf = 0;                    // 0: gradients by backpropagation
        std::vector<float> freqs, bins, w;
        int dtnyq = 1, ntnyq = 1, nq = 0;
        float df = 0;
        float *fact = nullptr;         // [ntnyq][nf] complex
        float *Ff = nullptr, *Fa = nullptr;   // [nq][nf][nfield] complex
        float *par = nullptr;          // [5][N]: M mu taup taus rho, unpadded
        float *w_dev = nullptr;        // [nf] angular frequencies
        size_t cap = 0;                // nf the buffers hold
    } dft;

    float *dalloc(long n) {
        void *p = nullptr;
        if (n <= 0) return nullptr;
        if (ctx->alloc((size_t)n * sizeof(float), &p, setting_up ? ctx->xfer : nullptr))
            return nullptr;
        owned.push_back(p);
        bytes += (size_t)n * sizeof(float);
        return (float *)p;
    }
    // n wavefield / memory-variable elements
    float *salloc(long n) { return dalloc((n * selem + 3) / 4); }
    // n coefficient or CPML-memory elements
    float *calloc_(long n) { return dalloc(n * (celem / 4)); }
    ~seis_model() {
        for (void *p : owned) ctx->release(p);
        void *sp[] = {shot.src_cell, shot.src_type, shot.src_k, shot.src, shot.rec_cell, shot.out,
                      grad.bnd, grad.res, grad.ckpt, grad.hist, grad.res_fac, sg_t};
        for (void *p : sp)
            if (p) ctx->release(p);
        for (auto e : evpool) (void)hipEventDestroy(e);
        if (sg_exec) (void)hipGraphExecDestroy(sg_exec);
        if (sg_graph) (void)hipGraphDestroy(sg_graph);
        fs_graph.reset();
        ad_graph.reset();
    }
};

// A batch of modeling jobs advanced together, one launch per pass and step
// for all of them (serving).  Slots hold models with a bound shot.
struct seis_batch {
    seis_ctx *ctx = nullptr;
    int max_slots = 0;
    std::vector<seis_model *> model;     // nullptr: free slot
    std::vector<int> t;                  // host copy of each slot's next step
    std::vector<int> tmax;
    BatchSlot *dslots = nullptr;         // device slot table
    int *dids = nullptr;                 // device slot table in class order
    unsigned *dtile0 = nullptr;
    BatchTable table{};
    int step = 0;                        // the next batch step to launch
    int maxrec = 0;
    bool dirty = true;
    // host sources of asynchronous copies
    std::vector<BatchSlot> host_slots;
    std::vector<int> host_active;
    // class tables, double buffered: a rebuild writes the buffer the queued
    // launches are not reading, after the launches that last read it ended
    std::vector<int> host_ids[2];
    std::vector<unsigned> host_tile0[2];
    hipEvent_t table_ev[2] = {nullptr, nullptr};
    int tbuf = 0;
    // completion points of the steps, reused round robin
    std::vector<hipEvent_t> step_ev;
    size_t step_ev_next = 0;
};

namespace {

// Small host float array (profiles) -> device array of the coefficient type.
int upload_raw(seis_model *m, const float **dst, const std::vector<float> &host) {
    float *d = m->calloc_((long)host.size());
    if (!d) return 1;
    hipStream_t us = m->setting_up ? m->ctx->xfer : m->ctx->stream;
    if (m->celem == 8) {
        std::vector<double> h(host.begin(), host.end());
        HIP_TRY(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, us));
        HIP_TRY(hipStreamSynchronize(us));
    } else {
        HIP_TRY(hipMemcpyAsync(d, host.data(), host.size() * 4, hipMemcpyHostToDevice, us));
        HIP_TRY(hipStreamSynchronize(us));
    }
    *dst = d;
    return 0;
}

// Unpadded reference array (z fastest) -> device coefficient array in the
// padded layout (float, or double in FP64 mode), optionally multiplied by a
// power of two (exact).  The host expansion runs on all host threads.
int upload(seis_model *m, const float **dst, const std::vector<float> &ref, float mul = 1) {
    float *d = m->calloc_(m->g.nfield);
    if (!d) return 1;
    hipStream_t us = m->setting_up ? m->ctx->xfer : m->ctx->stream;
    if (m->celem == 8) {
        std::vector<double> h;
        to_padded(m->g, ref.data(), mul, h);
        HIP_TRY(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, us));
        HIP_TRY(hipStreamSynchronize(us));
    } else {
        std::vector<float> h;
        to_padded(m->g, ref.data(), mul, h);
        HIP_TRY(hipMemcpyAsync(d, h.data(), h.size() * 4, hipMemcpyHostToDevice, us));
        HIP_TRY(hipStreamSynchronize(us));
    }
    *dst = d;
    return 0;
}

int out_code(const std::string &n) {
    static const std::map<std::string, int> k = {
        {"vx", O_VX},   {"vy", O_VY},   {"vz", O_VZ},   {"sxx", O_SXX}, {"syy", O_SYY},
        {"szz", O_SZZ}, {"sxz", O_SXZ}, {"sxy", O_SXY}, {"syz", O_SYZ}, {"p", O_P}};
    return k.at(n);
}

// Derived per-cell coefficients of the viscoelastic stress update, with the
// reference's expressions (src/update_s3D.cl:324-340, src/update_s2D.cl:252-259,
// src/update_s2D_acc.cl:343-346).
int upload_params(seis_model *m, const RefParams &p) {
    const seis_csts &c = m->c;
    const int L = c.L;
    const float DT = m->s.DT;
    FDArgs<float> &a = m->a;
    // scaled units: buoyancies times q, stress-update coefficients over q
    const float qv = m->q, qs = 1.0f / m->q;
    if (upload(m, &a.rip, p.rip, qv) || upload(m, &a.rkp, p.rkp, qv)) return 1;
    if (c.ND == 3 && upload(m, &a.rjp, p.rjp, qv)) return 1;
    if (L == 0) {
        if (upload(m, &a.M, p.M, qs)) return 1;
        if (c.ND != 22) {
            if (upload(m, &a.mu, p.mu, qs) || upload(m, &a.muipkp, p.muipkp, qs)) return 1;
            if (c.ND == 3 &&
                (upload(m, &a.muipjp, p.muipjp, qs) || upload(m, &a.mujpkp, p.mujpkp, qs)))
                return 1;
        }
        return 0;
    }
    std::vector<float> g, f, d, e;
    visco_coeffs(c, p, DT, g, f, d, e);
    if (c.ND == 22) return upload(m, &a.cg, g, qs) || upload(m, &a.ce, e, qs);
    // the adjoint free surface reads the relaxed moduli (src/surface2D_adj.cl:
    // 67-68, src/surface3D_adj.cl), which the viscoelastic update does not
    if (c.freesurf && (upload(m, &a.M, p.M, qs) || upload(m, &a.mu, p.mu, qs))) return 1;
    if (upload(m, &a.cg, g, qs) || upload(m, &a.cf, f, qs) || upload(m, &a.cd, d, qs) ||
        upload(m, &a.ce, e, qs))
        return 1;
    auto fij = [&](const std::vector<float> &mu, const std::vector<float> &ts, const float **fo,
                   const float **dO) {
        std::vector<float> ff, dd;
        visco_pair(mu, ts, L, DT, ff, dd);
        return upload(m, fo, ff, qs) || upload(m, dO, dd, qs);
    };
    if (fij(p.muipkp, p.tausipkp, &a.fipkp, &a.dipkp)) return 1;
    if (c.ND == 3) {
        if (fij(p.muipjp, p.tausipjp, &a.fipjp, &a.dipjp)) return 1;
        if (fij(p.mujpkp, p.tausjpkp, &a.fjpkp, &a.djpkp)) return 1;
    }
    return 0;
}

int alloc_fields(seis_model *m) {
    const seis_csts &c = m->c;
    FDArgs<float> &a = m->a;
    const long nf = m->g.nfield, L = c.L;
    a.vx = m->salloc(nf);
    a.vz = m->salloc(nf);
    if (c.ND == 3) a.vy = m->salloc(nf);
    if (c.ND == 22) {
        a.p = m->salloc(nf);
        if (L) a.rp = m->salloc(nf * L);
    } else {
        a.sxx = m->salloc(nf);
        a.szz = m->salloc(nf);
        a.sxz = m->salloc(nf);
        if (L) {
            a.rxx = m->salloc(nf * L);
            a.rzz = m->salloc(nf * L);
            a.rxz = m->salloc(nf * L);
        }
        if (c.ND == 3) {
            a.syy = m->salloc(nf);
            a.sxy = m->salloc(nf);
            a.syz = m->salloc(nf);
            if (L) {
                a.ryy = m->salloc(nf * L);
                a.rxy = m->salloc(nf * L);
                a.ryz = m->salloc(nf * L);
            }
        }
    }
    if (c.abs_type == 1) {
        const long nab2 = 2L * c.nab, NZ = m->g.NZ, NY = m->g.NY, NX = m->g.NX;
        const long sxs = nab2 * NY * NZ, sys = NX * nab2 * NZ, szs = NX * NY * nab2;
        if (c.ND == 22) {
            a.psi_p_x = m->calloc_(sxs);
            a.psi_p_z = m->calloc_(szs);
            a.psi_vx_x = m->calloc_(sxs);
            a.psi_vz_z = m->calloc_(szs);
        } else {
            a.psi_sxx_x = m->calloc_(sxs);
            a.psi_sxz_x = m->calloc_(sxs);
            a.psi_szz_z = m->calloc_(szs);
            a.psi_sxz_z = m->calloc_(szs);
            a.psi_vx_x = m->calloc_(sxs);
            a.psi_vz_x = m->calloc_(sxs);
            a.psi_vx_z = m->calloc_(szs);
            a.psi_vz_z = m->calloc_(szs);
            if (c.ND == 3) {
                a.psi_sxy_x = m->calloc_(sxs);
                a.psi_syz_z = m->calloc_(szs);
                a.psi_vy_x = m->calloc_(sxs);
                a.psi_vy_z = m->calloc_(szs);
                a.psi_syy_y = m->calloc_(sys);
                a.psi_sxy_y = m->calloc_(sys);
                a.psi_syz_y = m->calloc_(sys);
                a.psi_vx_y = m->calloc_(sys);
                a.psi_vy_y = m->calloc_(sys);
                a.psi_vz_y = m->calloc_(sys);
            }
        }
    }
    return g_err.empty() ? 0 : 1;
}

// Zero every wavefield, memory variable and CPML memory of the model
// (vars_init, src/automatic_kernels.c:291-392).
void zero_state(seis_model *m) {
    const FDArgs<float> &a = m->a;
    const size_t nf = ((size_t)m->g.nfield * m->selem + 3) / 4 * 4;
    const size_t L = m->c.L;
    float *f1[] = {a.vx, a.vy, a.vz, a.sxx, a.syy, a.szz, a.sxy, a.syz, a.sxz, a.p};
    for (float *p : f1) launch_fill(p, nf, m->ctx->stream);
    float *fl[] = {a.rxx, a.ryy, a.rzz, a.rxy, a.ryz, a.rxz, a.rp};
    for (float *p : fl) launch_fill(p, nf * L, m->ctx->stream);
    if (m->c.abs_type == 1) {
        const long nab2 = 2L * m->c.nab, NZ = m->g.NZ, NY = m->g.NY, NX = m->g.NX;
        const size_t ce = m->celem;
        const size_t sx = nab2 * NY * NZ * ce, sy = NX * nab2 * NZ * ce, sz = NX * NY * nab2 * ce;
        float *px[] = {a.psi_sxx_x, a.psi_sxy_x, a.psi_sxz_x, a.psi_vx_x, a.psi_vy_x, a.psi_vz_x, a.psi_p_x};
        float *py[] = {a.psi_sxy_y, a.psi_syy_y, a.psi_syz_y, a.psi_vx_y, a.psi_vy_y, a.psi_vz_y};
        float *pz[] = {a.psi_sxz_z, a.psi_syz_z, a.psi_szz_z, a.psi_vx_z, a.psi_vy_z, a.psi_vz_z, a.psi_p_z};
        for (float *p : px) launch_fill(p, sx, m->ctx->stream);
        for (float *p : py) launch_fill(p, sy, m->ctx->stream);
        for (float *p : pz) launch_fill(p, sz, m->ctx->stream);
    }
}

float *field_ptr(seis_model *m, const std::string &n) {
    const FDArgs<float> &a = m->a;
    if (n == "vx") return a.vx;
    if (n == "vy") return a.vy;
    if (n == "vz") return a.vz;
    if (n == "sxx") return a.sxx;
    if (n == "syy") return a.syy;
    if (n == "szz") return a.szz;
    if (n == "sxy") return a.sxy;
    if (n == "syz") return a.syz;
    if (n == "sxz") return a.sxz;
    if (n == "p") return a.p;
    return nullptr;
}

hipEvent_t pool_event(seis_model *m, size_t k) {
    while (m->evpool.size() <= k) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        m->evpool.push_back(e);
    }
    return m->evpool[k];
}

// IEEE binary16 -> float (host side of FP16 storage)
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31, ex = (h >> 10) & 0x1f, man = h & 0x3ff;
    float f;
    if (ex == 0) f = std::ldexp((float)man, -24);                    // zero, subnormal
    else if (ex == 31) f = man ? NAN : INFINITY;
    else f = std::ldexp((float)(man | 0x400), (int)ex - 25);
    uint32_t bits;
    std::memcpy(&bits, &f, 4);
    bits |= sign;
    std::memcpy(&f, &bits, 4);
    return f;
}

// The FP64 view of the model's buffers: the same device arrays (allocated
// with 8-byte elements) and the scalars promoted exactly.
FDArgs<double> as_double(const FDArgs<float> &a) {
    FDArgs<double> d;
    std::memset(&d, 0, sizeof(d));
    d.g = a.g;
    for (int k = 0; k < 7; ++k) d.hc[k] = a.hc[k];
    d.DT = a.DT;
    d.DT2 = a.DT2;
    d.L = a.L;
    for (int l = 0; l < kMaxL; ++l) {
        d.eta[l] = a.eta[l];
        d.bl[l] = a.bl[l];
        d.cl[l] = a.cl[l];
        d.bs[l] = a.bs[l];
    }
    d.abs_type = a.abs_type;
    d.freesurf = a.freesurf;
    d.nab = a.nab;
    d.fuse_surface = a.fuse_surface;
#define SEIS_AS_D(f) d.f = (decltype(d.f))(a.f)
    SEIS_AS_D(vx); SEIS_AS_D(vy); SEIS_AS_D(vz); SEIS_AS_D(sxx); SEIS_AS_D(syy); SEIS_AS_D(szz);
    SEIS_AS_D(sxy); SEIS_AS_D(syz); SEIS_AS_D(sxz); SEIS_AS_D(p);
    SEIS_AS_D(rxx); SEIS_AS_D(ryy); SEIS_AS_D(rzz); SEIS_AS_D(rxy); SEIS_AS_D(ryz); SEIS_AS_D(rxz);
    SEIS_AS_D(rp);
    SEIS_AS_D(rip); SEIS_AS_D(rjp); SEIS_AS_D(rkp); SEIS_AS_D(M); SEIS_AS_D(mu);
    SEIS_AS_D(muipkp); SEIS_AS_D(muipjp); SEIS_AS_D(mujpkp); SEIS_AS_D(cg); SEIS_AS_D(cf);
    SEIS_AS_D(cd); SEIS_AS_D(ce); SEIS_AS_D(fipkp); SEIS_AS_D(fipjp); SEIS_AS_D(fjpkp);
    SEIS_AS_D(dipkp); SEIS_AS_D(dipjp); SEIS_AS_D(djpkp); SEIS_AS_D(cpml); SEIS_AS_D(taper);
    SEIS_AS_D(psi_sxx_x); SEIS_AS_D(psi_sxy_x); SEIS_AS_D(psi_sxz_x); SEIS_AS_D(psi_sxy_y);
    SEIS_AS_D(psi_syy_y); SEIS_AS_D(psi_syz_y); SEIS_AS_D(psi_sxz_z); SEIS_AS_D(psi_syz_z);
    SEIS_AS_D(psi_szz_z); SEIS_AS_D(psi_vx_x); SEIS_AS_D(psi_vy_x); SEIS_AS_D(psi_vz_x);
    SEIS_AS_D(psi_vx_y); SEIS_AS_D(psi_vy_y); SEIS_AS_D(psi_vz_y); SEIS_AS_D(psi_vx_z);
    SEIS_AS_D(psi_vy_z); SEIS_AS_D(psi_vz_z); SEIS_AS_D(psi_p_x); SEIS_AS_D(psi_p_z);
#undef SEIS_AS_D
    return d;
}

// One launch of each step kernel in the model's precision.
// tb (may be NULL): device step counter added to t (step graphs)
void step_sources(seis_model *m, int t, int pdir, hipStream_t st, const int *tb = nullptr) {
    if (m->prec == P_FP16)
        launch_sources<float, __half>(m->a, m->nd_case, m->shot, m->c.NT, t, pdir, st, tb);
    else if (m->prec == P_FP64)
        launch_sources<double, double>(m->ad, m->nd_case, m->shot, m->c.NT, t, pdir, st, tb);
    else
        launch_sources<float, float>(m->a, m->nd_case, m->shot, m->c.NT, t, pdir, st, tb);
}
void step_v(seis_model *m, hipStream_t st) {
    if (m->prec == P_FP16) launch_update_v<float, __half>(m->a, m->nd_case, st);
    else if (m->prec == P_FP64) launch_update_v(m->ad, m->nd_case, st);
    else launch_update_v(m->a, m->nd_case, st);
}
void step_s(seis_model *m, hipStream_t st) {
    if (m->prec == P_FP16) launch_update_s<float, __half>(m->a, m->nd_case, st);
    else if (m->prec == P_FP64) launch_update_s(m->ad, m->nd_case, st);
    else launch_update_s(m->a, m->nd_case, st);
}
// the free surface as its own kernel (2D, and 3D runs that save boundaries)
void step_surface(seis_model *m, hipStream_t st) {
    if (m->prec == P_FP16) launch_surface<float, __half>(m->a, m->nd_case, 1, st);
    else if (m->prec == P_FP64) launch_surface(m->ad, m->nd_case, 1, st);
    else launch_surface(m->a, m->nd_case, 1, st);
}
void step_varsout(seis_model *m, int t, hipStream_t st, const int *tb = nullptr) {
    const int n = (int)m->out_codes.size();
    const float *u = m->unscale.empty() ? nullptr : m->unscale.data();
    if (m->prec == P_FP16)
        launch_varsout<float, __half>(m->a, m->nd_case, m->shot, m->out_codes.data(), u, n, t, st,
                                      tb);
    else if (m->prec == P_FP64)
        launch_varsout<double, double>(m->ad, m->nd_case, m->shot, m->out_codes.data(), u, n, t,
                                       st, tb);
    else
        launch_varsout<float, float>(m->a, m->nd_case, m->shot, m->out_codes.data(), u, n, t, st,
                                     tb);
}

// varsout of step t_rec and the sources of step t_src in one launch
void step_io(seis_model *m, int t_rec, int t_src, hipStream_t st, const int *tb = nullptr) {
    const int n = (int)m->out_codes.size(), NT = m->c.NT;
    const float *u = m->unscale.empty() ? nullptr : m->unscale.data();
    const int *oc = m->out_codes.data();
    if (m->prec == P_FP16)
        launch_io<float, __half>(m->a, m->nd_case, m->shot, oc, u, n, NT, t_rec, t_src, st, tb);
    else if (m->prec == P_FP64)
        launch_io<double, double>(m->ad, m->nd_case, m->shot, oc, u, n, NT, t_rec, t_src, st, tb);
    else
        launch_io<float, float>(m->a, m->nd_case, m->shot, oc, u, n, NT, t_rec, t_src, st, tb);
}

// One forward step t (src/time_stepping.c:785-816) led by varsout of step
// t-1, which shares the sources' launch: [varsout t-1, sources t], update_v,
// update_s, the free surface unless fused.  A run of steps ends with
// step_io(m, last, -1) for its last varsout.  Recording step t-1 again at
// the start of a run rewrites the same values.
void step_all(seis_model *m, int t, hipStream_t st, const int *tb = nullptr) {
    step_io(m, t - 1, t, st, tb);
    step_v(m, st);
    step_s(m, st);
    if (m->c.freesurf && !m->a.fuse_surface) step_surface(m, st);
}

// What a captured step sequence depends on: the kernel arguments (model
// arguments, shot tables, outputs and their scales) and the precision.
uint64_t step_key(const seis_model *m) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(&m->a, sizeof(m->a));
    mix(&m->ad, sizeof(m->ad));
    mix(&m->shot, sizeof(m->shot));
    mix(&m->prec, sizeof(m->prec));
    mix(m->out_codes.data(), m->out_codes.size() * sizeof(int));
    mix(m->unscale.data(), m->unscale.size() * sizeof(float));
    return h;
}

// Steps per captured graph of forward steps.  2D grids run each kernel in a
// few microseconds, so host launches, not the GPU, set the pace of a step;
// replaying a graph of kGraphSteps steps submits them at once.  The step
// index comes from a device counter that the graph's last node advances.
constexpr int kGraphSteps = 64;

bool graph_enabled() {
    static const bool on = [] {
        const char *v = getenv("SEIS_STEP_GRAPH");
        return !(v && *v == '0');
    }();
    return on;
}

void graph_reset(seis_model *m) {
    if (m->sg_exec) (void)hipGraphExecDestroy(m->sg_exec);
    if (m->sg_graph) (void)hipGraphDestroy(m->sg_graph);
    m->sg_exec = nullptr;
    m->sg_graph = nullptr;
    m->sg_key = 0;
}

// The device step counter of replayed step graphs.
int step_counter(seis_model *m) {
    if (m->sg_t) return 0;
    void *p = nullptr;
    if (m->ctx->alloc(256, &p)) return 1;
    m->sg_t = (int *)p;
    return 0;
}

// key mixed with n more bytes
uint64_t key_mix(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// Capture body() on stream st into G unless G already holds a graph of the
// same key.  body only enqueues kernels.
template <typename F>
int capture_steps(seis_model::Graph &G, uint64_t key, hipStream_t st, F &&body) {
    if (G.e && G.key == key) return 0;
    G.reset();
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return fail(std::string("step graph capture: ") + hipGetErrorString(e));
    }
    G.g = g;
    HIP_TRY(hipGraphInstantiate(&G.e, g, nullptr, nullptr, 0));
    G.key = key;
    return 0;
}

// Capture (or reuse) the graph of kGraphSteps steps for the model's current
// arguments on stream st.
int graph_prepare(seis_model *m, hipStream_t st) {
    const uint64_t key = step_key(m);
    if (m->sg_exec && m->sg_key == key) return 0;
    graph_reset(m);
    if (step_counter(m)) return 1;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    for (int k = 0; k < kGraphSteps; ++k) step_all(m, k, st, m->sg_t);
    launch_tcounter(m->sg_t, kGraphSteps, 1, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return fail(std::string("step graph capture: ") + hipGetErrorString(e));
    }
    m->sg_graph = g;
    HIP_TRY(hipGraphInstantiate(&m->sg_exec, g, nullptr, nullptr, 0));
    m->sg_key = key;
    return 0;
}

// Gradient state: adjoint wavefields, their CPML memory (or the forward
// run's, as the reference does, without FIX_ADJ_CPML), double accumulators.
int grad_setup(seis_model *m) {
    auto &G = m->grad;
    if (G.ready) return 0;
    const seis_csts &c = m->c;
    if (c.ND == 22)
        return fail("gradients need the elastic or viscoelastic case (ND 2 or 3)");
    const bool scaled = m->prec == P_SCALED || m->prec == P_FP16;
    if (scaled && c.L == 0)
        return fail("elastic gradients need FP32 or FP64 wavefields (FP16=0 or 64); "
                    "viscoelastic ones also run with FP16 storage (FP16=2) or scaled units "
                    "(FP16=1)");
    if (scaled && c.abs_type == 1 && !(c.gradfix & FIX_ADJ_CPML))
        return fail("gradients with scaled wavefields (FP16=1 or 2) need the adjoint run's own "
                    "CPML memory (gradfix bit 2)");
    G.visco = c.L > 0;
    if (!G.visco) {
        G.sh = make_shell(m->g, c.abs_type > 0 ? c.nab : 0, c.freesurf);
        if (G.sh.NXbnd < 2 * G.sh.F || G.sh.NZbnd < 2 * G.sh.F ||
            (c.ND == 3 && G.sh.NYbnd < 2 * G.sh.F))
            return fail("the grid inside the absorbing layers is too small for the boundary shell");
    }
    G.nvb = c.ND == 3 ? 9 : 5;
    FDArgs<float> &r = G.r;
    r = m->a;
    // adjoint fields, their memory and CPML memory are FP32 (FP64 in FP64 runs)
    const long ae = m->prec == P_FP64 ? 2 : 1;
    const long nf = m->g.nfield * ae;
    float **fl[] = {&r.vx, &r.vz, &r.sxx, &r.szz, &r.sxz, &r.vy, &r.syy, &r.sxy, &r.syz};
    const int nfl = c.ND == 3 ? 9 : 5;
    for (int q = 0; q < nfl; ++q)
        if (!(*fl[q] = m->dalloc(nf))) return 1;
    if (G.visco) {
        // the adjoint run's own memory variables, L slabs per component
        float **mv[] = {&r.rxx, &r.rzz, &r.rxz, &r.ryy, &r.rxy, &r.ryz};
        for (int q = 0; q < (c.ND == 3 ? 6 : 3); ++q)
            if (!(*mv[q] = m->dalloc(nf * c.L))) return 1;
        if (!(G.Ms = m->dalloc(nf)) || !(G.mus = m->dalloc(nf))) return 1;
    }
    // own adjoint CPML memory: zeroed (FIX_ADJ_CPML), or, for viscoelastic
    // runs whose forward memory is overwritten by the recomputation, a copy
    // of the forward run's final memory (what the reference's kernels bind)
    if (c.abs_type == 1 && ((c.gradfix & FIX_ADJ_CPML) || G.visco)) {
        const long nab2 = 2L * c.nab, NZ = m->g.NZ, NY = m->g.NY, NX = m->g.NX;
        const long sxs = nab2 * NY * NZ * ae, sys = NX * nab2 * NZ * ae, szs = NX * NY * nab2 * ae;
        float **px[] = {&r.psi_sxx_x, &r.psi_sxy_x, &r.psi_sxz_x, &r.psi_vx_x, &r.psi_vy_x,
                        &r.psi_vz_x};
        float **py[] = {&r.psi_sxy_y, &r.psi_syy_y, &r.psi_syz_y, &r.psi_vx_y, &r.psi_vy_y,
                        &r.psi_vz_y};
        float **pz[] = {&r.psi_sxz_z, &r.psi_syz_z, &r.psi_szz_z, &r.psi_vx_z, &r.psi_vy_z,
                        &r.psi_vz_z};
        for (float **p : px)
            if (*p && !(*p = m->dalloc(sxs))) return 1;
        for (float **p : py)
            if (*p && !(*p = m->dalloc(sys))) return 1;
        for (float **p : pz)
            if (*p && !(*p = m->dalloc(szs))) return 1;
    }
    if (c.freesurf && !(G.stmp = m->dalloc(3L * m->g.NX * m->g.NY * ae))) return 1;
    const long nfd = 2L * m->g.nfield;   // one double per cell
    double **gd[] = {&G.gM, &G.gmu, &G.grho};
    for (double **p : gd)
        if (!(*p = (double *)m->dalloc(nfd))) return 1;
    if (c.Hout)
        for (double **p : {&G.hM, &G.hmu})
            if (!(*p = (double *)m->dalloc(nfd))) return 1;
    // tau gradients of viscoelastic runs (param_type 3 reads no tau)
    if (G.visco && c.param_type <= 2)
        for (double **p : {&G.gtp, &G.gts})
            if (!(*p = (double *)m->dalloc(nfd))) return 1;
    if ((c.gradfix & FIX_AVG) && c.restype != 1) {
        const int nsh = c.ND == 3 ? 3 : 1;
        for (int q = 0; q < nsh; ++q)
            if (!(G.gsh[q] = (double *)m->dalloc(nfd))) return 1;
        if (c.ND == 3 && !(G.gry = (double *)m->dalloc(nfd))) return 1;
        if (!(G.grz = (double *)m->dalloc(nfd))) return 1;
    }
    G.ready = true;
    return 0;
}

// zero the adjoint wavefields and the adjoint run's own CPML memory
void grad_zero_adjoint(seis_model *m) {
    const FDArgs<float> &r = m->grad.r, &a = m->a;
    const size_t ae = m->prec == P_FP64 ? 8 : 4;   // bytes of an adjoint value
    const size_t nf = (size_t)m->g.nfield * ae;
    float *fl[] = {r.vx, r.vy, r.vz, r.sxx, r.syy, r.szz, r.sxy, r.syz, r.sxz};
    for (float *p : fl) launch_fill(p, nf, m->ctx->stream);
    float *mv[] = {r.rxx, r.ryy, r.rzz, r.rxy, r.ryz, r.rxz};
    if (m->grad.visco)
        for (float *p : mv)
            if (p) launch_fill(p, nf * m->c.L, m->ctx->stream);
    const bool fix = m->c.gradfix & FIX_ADJ_CPML;
    if (m->c.abs_type == 1 && (fix || m->grad.visco)) {
        const long nab2 = 2L * m->c.nab, NZ = m->g.NZ, NY = m->g.NY, NX = m->g.NX;
        const size_t sx = nab2 * NY * NZ * ae, sy = NX * nab2 * NZ * ae, sz = NX * NY * nab2 * ae;
        float *px[] = {r.psi_sxx_x, r.psi_sxy_x, r.psi_sxz_x, r.psi_vx_x, r.psi_vy_x, r.psi_vz_x};
        float *py[] = {r.psi_sxy_y, r.psi_syy_y, r.psi_syz_y, r.psi_vx_y, r.psi_vy_y, r.psi_vz_y};
        float *pz[] = {r.psi_sxz_z, r.psi_syz_z, r.psi_szz_z, r.psi_vx_z, r.psi_vy_z, r.psi_vz_z};
        float *fx[] = {a.psi_sxx_x, a.psi_sxy_x, a.psi_sxz_x, a.psi_vx_x, a.psi_vy_x, a.psi_vz_x};
        float *fy[] = {a.psi_sxy_y, a.psi_syy_y, a.psi_syz_y, a.psi_vx_y, a.psi_vy_y, a.psi_vz_y};
        float *fz[] = {a.psi_sxz_z, a.psi_syz_z, a.psi_szz_z, a.psi_vx_z, a.psi_vy_z, a.psi_vz_z};
        auto put = [&](float *dst, float *src, size_t bytes) {
            if (!dst) return;
            if (fix) launch_fill(dst, bytes, m->ctx->stream);
            else (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, m->ctx->stream);
        };
        for (int q = 0; q < 6; ++q) {
            put(px[q], fx[q], sx);
            put(py[q], fy[q], sy);
            put(pz[q], fz[q], sz);
        }
    }
}

// The model's current arguments with the adjoint run's wavefields and CPML
// memory swapped in (parameters may have changed since grad_setup).
FDArgs<float> adjoint_args(const seis_model *m) {
    FDArgs<float> r = m->a;
    const FDArgs<float> &q = m->grad.r;
    r.vx = q.vx; r.vy = q.vy; r.vz = q.vz;
    r.sxx = q.sxx; r.syy = q.syy; r.szz = q.szz;
    r.sxy = q.sxy; r.syz = q.syz; r.sxz = q.sxz;
    r.psi_sxx_x = q.psi_sxx_x; r.psi_sxy_x = q.psi_sxy_x; r.psi_sxz_x = q.psi_sxz_x;
    r.psi_sxy_y = q.psi_sxy_y; r.psi_syy_y = q.psi_syy_y; r.psi_syz_y = q.psi_syz_y;
    r.psi_sxz_z = q.psi_sxz_z; r.psi_syz_z = q.psi_syz_z; r.psi_szz_z = q.psi_szz_z;
    r.psi_vx_x = q.psi_vx_x; r.psi_vy_x = q.psi_vy_x; r.psi_vz_x = q.psi_vz_x;
    r.psi_vx_y = q.psi_vx_y; r.psi_vy_y = q.psi_vy_y; r.psi_vz_y = q.psi_vz_y;
    r.psi_vx_z = q.psi_vx_z; r.psi_vy_z = q.psi_vy_z; r.psi_vz_z = q.psi_vz_z;
    r.rxx = q.rxx; r.ryy = q.ryy; r.rzz = q.rzz; r.rxy = q.rxy; r.ryz = q.ryz; r.rxz = q.rxz;
    return r;
}

// The forward run's state (wavefields, memory variables, CPML memory) as
// (device pointer, bytes) pairs: what a checkpoint holds.
std::vector<std::pair<float *, size_t>> forward_state(const seis_model *m) {
    const FDArgs<float> &a = m->a;
    const size_t nf = (size_t)m->g.nfield * m->selem, L = m->c.L, ce = m->celem;
    std::vector<std::pair<float *, size_t>> st;
    float *fl[] = {a.vx, a.vy, a.vz, a.sxx, a.syy, a.szz, a.sxy, a.syz, a.sxz};
    for (float *p : fl)
        if (p) st.push_back({p, nf});
    float *mv[] = {a.rxx, a.ryy, a.rzz, a.rxy, a.ryz, a.rxz};
    for (float *p : mv)
        if (p) st.push_back({p, nf * L});
    if (m->c.abs_type == 1) {
        const size_t nab2 = 2L * m->c.nab, NZ = m->g.NZ, NY = m->g.NY, NX = m->g.NX;
        const size_t sx = nab2 * NY * NZ * ce, sy = NX * nab2 * NZ * ce, sz = NX * NY * nab2 * ce;
        float *px[] = {a.psi_sxx_x, a.psi_sxy_x, a.psi_sxz_x, a.psi_vx_x, a.psi_vy_x, a.psi_vz_x};
        float *py[] = {a.psi_sxy_y, a.psi_syy_y, a.psi_syz_y, a.psi_vx_y, a.psi_vy_y, a.psi_vz_y};
        float *pz[] = {a.psi_sxz_z, a.psi_syz_z, a.psi_szz_z, a.psi_vx_z, a.psi_vy_z, a.psi_vz_z};
        for (int q = 0; q < 6; ++q) {
            if (px[q]) st.push_back({px[q], sx});
            if (py[q]) st.push_back({py[q], sy});
            if (pz[q]) st.push_back({pz[q], sz});
        }
    }
    return st;
}

// the forward fields a viscoelastic adjoint step reads, in AdjArgs::st order
std::vector<float *> history_fields(const seis_model *m) {
    const FDArgs<float> &a = m->a;
    if (m->c.ND == 3) return {a.vx, a.vy, a.vz, a.sxx, a.syy, a.szz, a.sxy, a.syz, a.sxz};
    return {a.vx, a.vz, a.sxx, a.szz, a.sxz};
}

int tmax_of(const seis_model *m) {
    return m->c.tmax > 0 ? std::min(m->c.tmax, m->c.NT) : m->c.NT;
}

// receivers that share a cell must add their adjoint sources in order
int receivers_share_cells(const seis_model *m, const std::vector<int> &cells) {
    std::vector<int> v;
    for (int c : cells)
        if (c >= 0) v.push_back(c);
    std::sort(v.begin(), v.end());
    return std::adjacent_find(v.begin(), v.end()) != v.end();
}

}  // namespace

extern "C" {

const char *seis_last_error(void) { return g_err.c_str(); }

size_t seis_csts_size(void) { return sizeof(seis_csts); }

void seis_csts_default(seis_csts *c) {
    std::memset(c, 0, sizeof(*c));
    c->ND = 2;
    c->dh = 10;
    c->dt = 0.0008f;
    c->NT = 875;
    c->FDORDER = 8;
    c->MAXRELERROR = 1;
    c->abs_type = 1;
    c->nab = 16;
    c->abpc = 6;
    c->VPPML = 3500;
    c->NPOWER = 2;
    c->FPML = 15;
    c->K_MAX_CPML = 2;
    c->f0 = 15;
    c->FL[0] = 15;
    c->seisout = 2;
    c->gradfix = FIX_RES_SCALE | FIX_ADJ_CPML | FIX_RHO_SCALE | FIX_SURF_SYY | FIX_VISCO_F0 |
                 FIX_RMSNORM | FIX_HESS;
}

int seis_ctx_create(int device_ordinal, seis_ctx **out) {
    g_err.clear();
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device_ordinal < 0 || device_ordinal >= n)
        return fail("seis_ctx_create: no HIP device " + std::to_string(device_ordinal));
    auto ctx = std::make_unique<seis_ctx>();
    ctx->device = device_ordinal;
    HIP_TRY(hipSetDevice(device_ordinal));
    if (scratch_address(device_ordinal, &ctx->scratch)) return 1;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->xfer, hipStreamNonBlocking));
    *out = ctx.release();
    return 0;
}

int seis_ctx_destroy(seis_ctx *ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->xfer);
    for (auto &kv : ctx->caps) (void)hipFree(kv.first);
    for (auto &kv : ctx->cache) (void)hipEventDestroy(kv.second.second);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->xfer);
    delete ctx;
    return 0;
}

int seis_ctx_sync(seis_ctx *ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

size_t seis_ctx_footprint(const seis_ctx *ctx) { return ctx ? ctx->bytes : 0; }

int seis_buf_create(seis_ctx *ctx, size_t bytes, void **dptr) { return ctx->alloc(bytes, dptr); }
int seis_buf_free(seis_ctx *ctx, void *dptr) { return ctx->release(dptr); }
int seis_buf_send(seis_ctx *ctx, void *dptr, const void *host, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(dptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}
int seis_buf_read(seis_ctx *ctx, void *host, const void *dptr, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int seis_model_create(seis_ctx *ctx, const seis_csts *cin, const float *p0, const float *p1,
                      const float *rho, const float *taup, const float *taus, seis_model **out) {
    g_err.clear();
    *out = nullptr;
    if (!ctx || !cin) return fail("seis_model_create: null argument");
    const seis_csts c = *cin;
    if (c.ND != 2 && c.ND != 3 && c.ND != 22) return fail("ND must be 2, 3 or 22");
    if (c.L < 0 || c.L > 4) return fail("L must be in [0, 4]");
    if (c.FP16 != P_FP32 && c.FP16 != P_SCALED && c.FP16 != P_FP16 && c.FP16 != P_FP64)
        return fail("FP16 must be 0 (FP32), 1 (FP32, scaled units), 2 (FP16 storage) or 64 (FP64)");
    if (c.NT <= 0) return fail("NT must be positive");
    const int nd = c.ND == 3 ? 3 : 2;
    for (int d = 0; d < nd; ++d) {
        if (c.N[d] <= 0) return fail("grid sizes must be positive");
        if (c.abs_type > 0 && c.N[d] < 2 * c.nab)
            return fail("each grid dimension must hold both absorbing layers (N >= 2*nab)");
    }
    if (c.abs_type > 0 && (c.nab < 1 || c.nab > kMaxNab)) return fail("nab out of range");
    if (c.abs_type == 2 && c.nab < 2) return fail("the Cerjan taper needs nab >= 2");
    if (c.ND == 22 && c.freesurf) return fail("the free surface is not available for ND=22");
    if (out_names(c).empty()) return fail("seisout has no outputs for this ND");
    auto m = std::make_unique<seis_model>();
    m->ctx = ctx;
    m->c = c;
    // set-up traffic goes to the transfer stream, so a model can be built on
    // another host thread while the compute stream runs other jobs
    m->setting_up = true;
    m->nd_case = c.ND;
    std::string err;
    if (build_setup(c, m->s, err)) return fail(err);
    RefParams rp;
    if (transform_params(c, p0, p1, rho, taup, taus, rp, err)) return fail(err);
    float dts = 0;
    if (check_stability(c, rp, &dts)) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "Time step too large, to be stable, set dt<%f", dts);
        return fail(buf);
    }
    const int NZ = c.N[0], NY = c.ND == 3 ? c.N[1] : 1, NX = c.ND == 3 ? c.N[2] : c.N[1];
    m->prec = c.FP16;
    m->selem = c.FP16 == P_FP16 ? 2 : c.FP16 == P_FP64 ? 8 : 4;
    m->celem = c.FP16 == P_FP64 ? 8 : 4;
    m->g = make_layout(nd, NZ, NY, NX, c.FDORDER / 2, m->selem);
    // kernels address every array with 32-bit byte offsets
    if ((unsigned long long)m->g.nfield * std::max(m->selem, m->celem) >= (1ull << 32))
        return fail("grid too large: one padded field must stay below 4 GiB");
    if (c.FP16 == P_SCALED || c.FP16 == P_FP16) {
        // q = av/as ~ the largest impedance rho*vp = sqrt(M/buoyancy), as a
        // power of two so that rescaling the coefficients is exact
        double zmax = 0;
        for (size_t i = 0; i < rp.M.size(); ++i)
            if (rp.rip[i] > 0) zmax = std::max(zmax, std::sqrt((double)rp.M[i] / rp.rip[i]));
        m->q = zmax > 0 ? std::ldexp(1.0f, (int)std::lround(std::log2(zmax))) : 1.0f;
    }
    FDArgs<float> &a = m->a;
    std::memset(&a, 0, sizeof(a));
    a.g = m->g;
    for (int k = 0; k < 7; ++k) a.hc[k] = m->s.hc[k];
    a.DT = m->s.DT;
    a.DT2 = m->s.DT2;
    a.L = c.L;
    for (int l = 0; l < c.L; ++l) {
        a.eta[l] = m->s.eta[l];
        a.bl[l] = m->s.bl[l];
        a.cl[l] = m->s.cl[l];
        a.bs[l] = m->s.bs[l];
    }
    a.abs_type = c.abs_type;
    a.freesurf = c.freesurf;
    a.nab = c.nab;
    // forward modeling applies the 3D free surface inside the stress kernel
    a.fuse_surface = c.ND == 3 ? 1 : 0;
    HIP_TRY(hipSetDevice(ctx->device));
    if (upload_params(m.get(), rp)) return 1;
    if (c.abs_type == 1) {
        if (upload_raw(m.get(), &a.cpml, m->s.cpml)) return 1;
    } else if (c.abs_type == 2) {
        if (upload_raw(m.get(), &a.taper, m->s.taper)) return 1;
    }
    if (alloc_fields(m.get())) return 1;
    if (m->prec == P_FP64) m->ad = as_double(m->a);
    m->rp = std::move(rp);
    m->outn = out_names(c);
    for (auto &n : m->outn) m->out_codes.push_back(out_code(n));
    HIP_TRY(hipStreamSynchronize(ctx->xfer));
    m->setting_up = false;
    *out = m.release();
    return 0;
}

int seis_model_destroy(seis_model *m) {
    if (!m) return 0;
    // buffers return to the context's cache behind a compute-stream event
    (void)hipStreamSynchronize(m->ctx->xfer);
    delete m;
    return 0;
}

int seis_model_set_params(seis_model *m, const float *p0, const float *p1, const float *rho,
                          const float *taup, const float *taus) {
    g_err.clear();
    std::string err;
    RefParams rp;
    if (transform_params(m->c, p0, p1, rho, taup, taus, rp, err)) return fail(err);
    float dts = 0;
    if (check_stability(m->c, rp, &dts)) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "Time step too large, to be stable, set dt<%f", dts);
        return fail(buf);
    }
    // free the previous coefficient arrays and upload the new ones
    FDArgs<float> &a = m->a;
    const float **pp[] = {&a.rip, &a.rjp, &a.rkp, &a.M, &a.mu, &a.muipkp, &a.muipjp, &a.mujpkp,
                          &a.cg, &a.cf, &a.cd, &a.ce, &a.fipkp, &a.fipjp, &a.fjpkp, &a.dipkp,
                          &a.dipjp, &a.djpkp};
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    for (const float **p : pp) {
        if (!*p) continue;
        auto it = std::find(m->owned.begin(), m->owned.end(), (void *)*p);
        if (it != m->owned.end()) m->owned.erase(it);
        m->bytes -= (size_t)m->g.nfield * m->celem;
        m->ctx->release((void *)*p);
        *p = nullptr;
    }
    if (upload_params(m, rp)) return 1;
    if (m->prec == P_FP64) m->ad = as_double(m->a);
    m->rp = std::move(rp);
    return 0;
}

int seis_model_set_params_device(seis_model *m, const float *dp0, const float *dp1,
                                 const float *drho, const float *dtaup, const float *dtaus) {
    g_err.clear();
    const size_t n = (size_t)m->g.NZ * m->g.NY * m->g.NX;
    const float *src[5] = {dp0, dp1, drho, dtaup, dtaus};
    std::vector<float> h[5];
    const float *hp[5] = {};
    for (int k = 0; k < 5; ++k) {
        if (!src[k]) continue;
        h[k].resize(n);
        HIP_TRY(hipMemcpyAsync(h[k].data(), src[k], n * 4, hipMemcpyDeviceToHost, m->ctx->xfer));
        hp[k] = h[k].data();
    }
    HIP_TRY(hipStreamSynchronize(m->ctx->xfer));
    return seis_model_set_params(m, hp[0], hp[1], hp[2], hp[3], hp[4]);
}

size_t seis_model_footprint(const seis_model *m) { return m ? m->bytes : 0; }

int seis_out_count(const seis_model *m) { return (int)m->outn.size(); }
const char *seis_out_name(const seis_model *m, int i) {
    if (i < 0 || i >= (int)m->outn.size()) return nullptr;
    return m->outn[i].c_str();
}

namespace {
// Host side of a shot: source and receiver cells, the source samples in the
// model's units (FP16 storage and scaled units fix av and as per shot).
struct ShotHost {
    std::vector<int> scell, stype, sk, rcell;
    std::vector<float> ss;      // scaled source samples (empty: src as given)
};

static int shot_host(seis_model *m, int nsrc, const float *src_pos, const float *src, int nrec,
                     const float *rec_pos, ShotHost &h) {
    const seis_csts &c = m->c;
    const Layout &g = m->g;
    const int H = g.H, NT = c.NT;
    const float DH = m->s.DH;
    h.scell.assign(nsrc, 0);
    h.stype.assign(nsrc, 0);
    h.sk.assign(nsrc, 0);
    h.rcell.assign(nrec, 0);
    h.ss.clear();
    auto cell = [&](const float *p, int &k, int &out) -> int {
        const int i = (int)(p[0] / DH) + H, j = (int)(p[1] / DH) + H;
        k = (int)(p[2] / DH) + H;
        if (k < 0 || k >= g.NZ + 2 * H || (c.ND == 3 && (j < 0 || j >= g.NY + 2 * H)))
            return fail("source or receiver outside the padded grid");
        if (i < H || i > g.NX + H - 1) {
            out = -1;   // outside this domain in x: skipped, as the reference does
            return 0;
        }
        out = (int)g.idx(k - H, c.ND == 3 ? j - H : 0, i - H);
        return 0;
    };
    for (int s = 0; s < nsrc; ++s) {
        if (cell(&src_pos[5 * s], h.sk[s], h.scell[s])) return 1;
        h.stype[s] = (int)src_pos[5 * s + 4];
        if (c.ND == 22 && h.stype[s] == 100)
            return fail("explosive sources (type 100) are not defined for ND=22; use type 3");
    }
    for (int r = 0; r < nrec; ++r) {
        int k;
        if (cell(&rec_pos[8 * r], k, h.rcell[r])) return 1;
    }
    if (nsrc && (m->prec == P_SCALED || m->prec == P_FP16)) {
        // scaled units: the largest per-step injection (DT * |src|) of the
        // sources feeding velocities (or, without any, stresses) is stored
        // as about 2^8; powers of two keep the scaling exact
        double af = 0, ae = 0;
        for (int k = 0; k < nsrc; ++k)
            for (int t = 0; t < NT; ++t) {
                const double v = std::fabs((double)src[(size_t)k * NT + t]) * c.dt;
                if (h.stype[k] <= 2) af = std::max(af, v);
                else ae = std::max(ae, v);
            }
        auto pow2 = [](double x) { return std::ldexp(1.0f, (int)std::lround(std::log2(x))); };
        if (af > 0) {
            m->av = pow2(256.0 / af);
            m->as = m->av / m->q;
        } else if (ae > 0) {
            m->as = pow2(256.0 / ae);
            m->av = m->as * m->q;
        } else {
            m->av = m->q;
            m->as = 1;
        }
        h.ss.resize((size_t)nsrc * NT);
        for (int k = 0; k < nsrc; ++k) {
            const float f = h.stype[k] <= 2 ? m->av : m->as;
            for (int t = 0; t < NT; ++t) h.ss[(size_t)k * NT + t] = src[(size_t)k * NT + t] * f;
        }
        m->unscale.clear();
        for (int code : m->out_codes)
            m->unscale.push_back(code <= O_VZ ? 1.0f / m->av : 1.0f / m->as);
    }
    return 0;
}

// Copy a prepared shot into the device buffers of sd (sized for it) on
// stream st (the transfer stream by default) and wait for the copies.
static int shot_upload(seis_model *m, const ShotHost &h, const float *src, ShotDev &sd,
                       hipStream_t st = nullptr) {
    const int nsrc = (int)h.scell.size(), nrec = (int)h.rcell.size(), NT = m->c.NT;
    if (!st) st = m->ctx->xfer;
    sd.nsrc = nsrc;
    sd.nrec = nrec;
    if (nsrc) {
        HIP_TRY(hipMemcpyAsync(sd.src_cell, h.scell.data(), nsrc * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(sd.src_type, h.stype.data(), nsrc * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(sd.src_k, h.sk.data(), nsrc * 4, hipMemcpyHostToDevice, st));
        const float *s = h.ss.empty() ? src : h.ss.data();
        HIP_TRY(hipMemcpyAsync(sd.src, s, (size_t)nsrc * NT * 4, hipMemcpyHostToDevice, st));
    }
    if (nrec)
        HIP_TRY(hipMemcpyAsync(sd.rec_cell, h.rcell.data(), nrec * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}
}  // namespace

int seis_shot_bind(seis_model *m, int nsrc, const float *src_pos, const float *src, int nrec,
                   const float *rec_pos) {
    g_err.clear();
    const int NT = m->c.NT;
    ShotHost h;
    if (shot_host(m, nsrc, src_pos, src, nrec, rec_pos, h)) return 1;
    seis_ctx *ctx = m->ctx;
    ShotDev &sd = m->shot;
    const int nout = (int)m->outn.size();
    if (nsrc > m->shot_nsrc_cap || nrec > m->shot_nrec_cap || !sd.out) {
        void *olds[] = {sd.src_cell, sd.src_type, sd.src_k, sd.src, sd.rec_cell, sd.out};
        for (void *p : olds) ctx->release(p);
        const int ns = std::max(nsrc, 1), nr = std::max(nrec, 1);
        hipStream_t xf = ctx->xfer;
        void *p;
        if (ctx->alloc(ns * 4, &p, xf)) return 1;
        sd.src_cell = (int *)p;
        if (ctx->alloc(ns * 4, &p, xf)) return 1;
        sd.src_type = (int *)p;
        if (ctx->alloc(ns * 4, &p, xf)) return 1;
        sd.src_k = (int *)p;
        if (ctx->alloc((size_t)ns * NT * 4, &p, xf)) return 1;
        sd.src = (float *)p;
        if (ctx->alloc(nr * 4, &p, xf)) return 1;
        sd.rec_cell = (int *)p;
        // FP64 runs record their seismograms in double
        if (ctx->alloc((size_t)nr * NT * nout * (m->prec == P_FP64 ? 8 : 4), &p, xf)) return 1;
        sd.out = (float *)p;
        m->shot_nsrc_cap = ns;
        m->shot_nrec_cap = nr;
    }
    if (shot_upload(m, h, src, sd)) return 1;
    m->rcell_host = h.rcell;
    m->recpos_host.assign(rec_pos, rec_pos + (size_t)8 * nrec);
    m->shot_bound = true;
    m->grad.saved = false;
    return 0;
}

int seis_launch_init(seis_model *m) {
    zero_state(m);
    if (m->shot.out)
        launch_fill(m->shot.out,
                    (size_t)m->shot.nrec * m->c.NT * m->outn.size() * (m->prec == P_FP64 ? 8 : 4),
                    m->ctx->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int seis_launch_sources(seis_model *m, int nt, int pdir) {
    if (!m->shot_bound) return fail("no shot bound");
    step_sources(m, nt, pdir, m->ctx->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}
int seis_launch_update_v(seis_model *m) {
    step_v(m, m->ctx->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}
int seis_launch_update_s(seis_model *m) {
    step_s(m, m->ctx->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}
int seis_launch_surface(seis_model *m) {
    // with the surface fused into update_s there is nothing left to launch
    if (!m->c.freesurf || m->a.fuse_surface) return 0;
    step_surface(m, m->ctx->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}
int seis_launch_varsout(seis_model *m, int nt) {
    if (!m->shot_bound) return fail("no shot bound");
    step_varsout(m, nt, m->ctx->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

namespace {
// The bound shot's seismograms as [nout][nrec][NT] into outf (float) or
// outd (double) (device: [NT][nout][nrec] of float, or of double in FP64 runs).
static int shot_read(seis_model *m, float *outf, double *outd) {
    const int NT = m->c.NT, nrec = m->shot.nrec, nout = (int)m->outn.size();
    const bool dbl = m->prec == P_FP64;
    std::vector<char> raw((size_t)NT * nout * nrec * (dbl ? 8 : 4));
    // a batch job is read on the transfer stream once its last step is
    // done, while the batch keeps running on the compute stream
    hipStream_t st = m->ctx->stream;
    if (m->done_ev) {
        st = m->ctx->xfer;
        HIP_TRY(hipStreamWaitEvent(st, m->done_ev, 0));
    }
    if (!raw.empty())
        HIP_TRY(hipMemcpyAsync(raw.data(), m->shot.out, raw.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const float *tf = (const float *)raw.data();
    const double *td = (const double *)raw.data();
    for (int o = 0; o < nout; ++o)
        for (int r = 0; r < nrec; ++r)
            for (int t = 0; t < NT; ++t) {
                const size_t k = ((size_t)t * nout + o) * nrec + r;
                const size_t j = ((size_t)o * nrec + r) * NT + t;
                if (outd) outd[j] = dbl ? td[k] : (double)tf[k];
                else outf[j] = dbl ? (float)td[k] : tf[k];
            }
    return 0;
}
}  // namespace

int seis_shot_read(seis_model *m, float *out) { return shot_read(m, out, nullptr); }

int seis_shot_read_f64(seis_model *m, double *out) {
    g_err.clear();
    if (m->prec != P_FP64) return fail("seis_shot_read_f64: the model does not run in FP64 (FP16=64)");
    return shot_read(m, nullptr, out);
}

int seis_run_steps(seis_model *m, int t0, int nsteps) {
    g_err.clear();
    if (!m->shot_bound) return fail("no shot bound");
    if (t0 < 0 || nsteps < 0 || t0 + nsteps > m->c.NT) return fail("steps outside [0, NT)");
    hipStream_t st = m->ctx->stream;
    hipEvent_t e0 = pool_event(m, 0), e1 = pool_event(m, 1);
    if (!e0 || !e1) return fail("hipEventCreate failed");
    const bool prof = m->profiling;
    if (prof && !pool_event(m, 2 + 4 * (size_t)nsteps)) return fail("hipEventCreate failed");
    // 2D runs without per-kernel events replay captured step graphs
    const bool graph = !prof && m->c.ND != 3 && nsteps >= kGraphSteps && graph_enabled();
    if (graph && graph_prepare(m, st)) return 1;
    HIP_TRY(hipEventRecord(e0, st));
    int k0 = 0;
    if (graph) {
        launch_tcounter(m->sg_t, t0, 0, st);
        for (; k0 + kGraphSteps <= nsteps; k0 += kGraphSteps) HIP_TRY(hipGraphLaunch(m->sg_exec, st));
    }
    for (int k = k0; k < nsteps; ++k) {
        const int t = t0 + k;
        if (!prof) {
            step_all(m, t, st);
            continue;
        }
        // step order of src/time_stepping.c:785-816
        step_sources(m, t, 1, st);
        if (prof) HIP_TRY(hipEventRecord(m->evpool[2 + 4 * k], st));
        step_v(m, st);
        if (prof) HIP_TRY(hipEventRecord(m->evpool[3 + 4 * k], st));
        if (prof) HIP_TRY(hipEventRecord(m->evpool[4 + 4 * k], st));
        step_s(m, st);
        if (prof) HIP_TRY(hipEventRecord(m->evpool[5 + 4 * k], st));
        if (m->c.freesurf && !m->a.fuse_surface) step_surface(m, st);
        step_varsout(m, t, st);
    }
    // the last step's varsout (recorded at the start of each step otherwise)
    if (!prof && nsteps > 0) step_io(m, t0 + nsteps - 1, -1, st);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    m->ms_loop = ms;
    m->ms_v = m->ms_s = 0;
    if (prof) {
        for (int k = 0; k < nsteps; ++k) {
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, m->evpool[2 + 4 * k], m->evpool[3 + 4 * k]));
            HIP_TRY(hipEventElapsedTime(&b, m->evpool[4 + 4 * k], m->evpool[5 + 4 * k]));
            m->ms_v += a;
            m->ms_s += b;
        }
    }
    return 0;
}

int seis_forward_shot(seis_model *m, int nsrc, const float *src_pos, const float *src, int nrec,
                      const float *rec_pos, float *out) {
    g_err.clear();
    if (seis_shot_bind(m, nsrc, src_pos, src, nrec, rec_pos)) return 1;
    if (seis_launch_init(m)) return 1;
    const int tmax = m->c.tmax > 0 ? std::min(m->c.tmax, m->c.NT) : m->c.NT;
    if (seis_run_steps(m, 0, tmax)) return 1;
    return out ? seis_shot_read(m, out) : 0;
}

int seis_forward_shot_static(seis_model *m, int nsrc, const float *src_pos, const float *src,
                             int nrec, const float *rec_pos, float *out) {
    g_err.clear();
    if (!m) return fail("seis_forward_shot_static: no model");
    if (m->in_batch) return fail("seis_forward_shot_static: the model is held by a serving batch");
    if (nsrc < 0 || nrec < 0) return fail("seis_forward_shot_static: negative source or receiver count");
    void *base = m->ctx->scratch;
    if (!base) return fail("seis_forward_shot_static: no static area on this device");
    ShotHost h;
    if (shot_host(m, nsrc, src_pos, src, nrec, rec_pos, h)) return 1;
    // carve the shot's device tables out of the area, 256-byte aligned
    const int NT = m->c.NT, nout = (int)m->outn.size();
    const int ns = std::max(nsrc, 1), nr = std::max(nrec, 1);
    const size_t elem = m->prec == P_FP64 ? 8 : 4;
    const size_t nseis = (size_t)nr * NT * nout;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_cell = carve(ns * 4), o_type = carve(ns * 4), o_k = carve(ns * 4);
    const size_t o_src = carve((size_t)ns * NT * 4), o_rec = carve(nr * 4);
    const size_t o_out = carve(nseis * elem), o_tr = carve(nseis * 4);
    if (off > SEIS_STATIC_SHOT_BYTES)
        return fail("seis_forward_shot_static: the shot needs " + std::to_string(off) +
                    " bytes of device memory, the static area holds " +
                    std::to_string((size_t)SEIS_STATIC_SHOT_BYTES));
    char *b = (char *)base;
    ShotDev sd;
    sd.src_cell = (int *)(b + o_cell);
    sd.src_type = (int *)(b + o_type);
    sd.src_k = (int *)(b + o_k);
    sd.src = (float *)(b + o_src);
    sd.rec_cell = (int *)(b + o_rec);
    sd.out = (float *)(b + o_out);
    float *tr = (float *)(b + o_tr);
    hipStream_t st = m->ctx->stream;
    // the model's own shot buffers are set aside for this run and kept
    const ShotDev own = m->shot;
    m->shot = sd;
    m->shot_bound = false;
    m->grad.saved = false;
    int rc = shot_upload(m, h, src, m->shot, st);
    if (!rc) {
        zero_state(m);
        launch_fill(sd.out, nseis * elem, st);
        const int tmax = m->c.tmax > 0 ? std::min(m->c.tmax, NT) : NT;
        for (int t = 0; t < tmax; ++t) {
            // step order of src/time_stepping.c:785-816, as seis_run_steps
            step_sources(m, t, 1, st);
            step_v(m, st);
            step_s(m, st);
            if (m->c.freesurf && !m->a.fuse_surface) step_surface(m, st);
            step_varsout(m, t, st);
        }
        const long n = (long)nseis;
        const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
        if (n > 0 && nrec > 0) {
            if (m->prec == P_FP64)
                hipLaunchKernelGGL(k_seis_transpose<double>, dim3(blocks), dim3(256), 0, st,
                                   (const double *)sd.out, tr, NT, nout, nrec);
            else
                hipLaunchKernelGGL(k_seis_transpose<float>, dim3(blocks), dim3(256), 0, st,
                                   (const float *)sd.out, tr, NT, nout, nrec);
        }
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && out && nrec > 0 && nout > 0)
            e = hipMemcpyAsync(out, tr, (size_t)nrec * NT * nout * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            rc = fail(std::string("seis_forward_shot_static: ") + hipGetErrorString(e));
    }
    m->shot = own;
    m->shot_bound = false;
    return rc;
}

int seis_model_set_profiling(seis_model *m, int on) {
    m->profiling = on != 0;
    return 0;
}

int seis_read_field(seis_model *m, const char *name, float *host) {
    g_err.clear();
    const std::string nm = name ? name : "";
    float *d = field_ptr(m, nm);
    if (!d) return fail(std::string("no wavefield named ") + (name ? name : "(null)"));
    const size_t nf = (size_t)m->g.nfield;
    std::vector<unsigned char> raw(nf * m->selem);
    HIP_TRY(hipMemcpyAsync(raw.data(), d, raw.size(), hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    std::vector<float> tmp(nf);
    const float u = (m->prec == P_SCALED || m->prec == P_FP16)
                        ? (nm[0] == 'v' ? 1.0f / m->av : 1.0f / m->as)
                        : 1.0f;
    for (size_t i = 0; i < nf; ++i) {
        if (m->selem == 8) tmp[i] = (float)((const double *)raw.data())[i];
        else if (m->selem == 2) tmp[i] = half_to_float(((const uint16_t *)raw.data())[i]) * u;
        else tmp[i] = ((const float *)raw.data())[i] * u;
    }
    const Layout &g = m->g;
    const int H = g.H;
    const long NZp = g.NZ + 2 * H, NYp = g.nd == 3 ? g.NY + 2 * H : 1, NXp = g.NX + 2 * H;
    for (long x = 0; x < NXp; ++x)
        for (long y = 0; y < NYp; ++y)
            for (long z = 0; z < NZp; ++z)
                host[(x * NYp + y) * NZp + z] =
                    tmp[g.idx((int)z - H, g.nd == 3 ? (int)y - H : 0, (int)x - H)];
    return 0;
}

// ---- gradient by backpropagation ------------------------------------------

int seis_grad_set_memory(seis_model *m, size_t bytes) {
    m->grad.mem_limit = bytes;
    return 0;
}

int seis_grad_plan(const seis_model *m, int *levels, int *K, int *n, int *n0, size_t *bytes) {
    const CkptPlan &p = m->grad.plan;
    if (levels) *levels = p.L;
    if (K)
        for (int l = 0; l < kMaxLevels; ++l) K[l] = l < p.L ? p.K[l] : 0;
    if (n) *n = p.n;
    if (n0) *n0 = p.n0;
    if (bytes) *bytes = p.bytes;
    return 0;
}

int seis_grad_zero(seis_model *m) {
    g_err.clear();
    if (grad_setup(m)) return 1;
    const size_t nb = (size_t)m->g.nfield * 8;
    launch_fill(m->grad.gM, nb, m->ctx->stream);
    launch_fill(m->grad.gmu, nb, m->ctx->stream);
    launch_fill(m->grad.grho, nb, m->ctx->stream);
    if (m->grad.hM) launch_fill(m->grad.hM, nb, m->ctx->stream);
    if (m->grad.hmu) launch_fill(m->grad.hmu, nb, m->ctx->stream);
    if (m->grad.gtp) launch_fill(m->grad.gtp, nb, m->ctx->stream);
    if (m->grad.gts) launch_fill(m->grad.gts, nb, m->ctx->stream);
    for (double *p : {m->grad.gsh[0], m->grad.gsh[1], m->grad.gsh[2], m->grad.gry, m->grad.grz})
        if (p) launch_fill(p, nb, m->ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return 0;
}

namespace {

// (Re)allocate a gradient work buffer of at least `bytes` (exact: of exactly
// `bytes`, freeing it when 0).
static int grad_buffer(seis_model *m, float **buf, size_t *have, size_t bytes, bool exact) {
    if (exact ? *have == bytes : *have >= bytes) return 0;
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    if (*buf) {
        m->bytes -= *have;
        m->ctx->release(*buf);
        *buf = nullptr;
        *have = 0;
    }
    if (!bytes) return 0;
    void *p = nullptr;
    if (m->ctx->alloc(bytes, &p)) return 1;
    *buf = (float *)p;
    *have = bytes;
    m->bytes += bytes;
    return 0;
}

// Checkpoint plan of a viscoelastic gradient over T steps, with L levels.
// The forward run keeps n0 = ceil(T / K[0]) checkpoints of the full state
// (S bytes each), one every K[0] steps.  The adjoint treats the level-0
// segments from the last to the first: it restores the segment's
// checkpoint and hands it to level 1, which recomputes it once to keep a
// checkpoint every K[1] = K[0] / n steps (n - 1 of them: the last
// sub-segment starts from the state the recomputation ends on), and so on
// down to level L-1, whose segments of K[L-1] steps are recomputed keeping
// every post-source state (H bytes each) for the adjoint steps that run
// backwards over them.  Memory: n0 S + (L-1)(n-1) S + K[L-1] H; work: one
// forward run plus L recomputations of it.  The plan has the fewest levels
// that fit the budget B, and of those the least memory.
static bool plan_checkpoints(int T, size_t S, size_t H, size_t B, CkptPlan &best) {
    best = CkptPlan{};
    for (int L = 1; L <= kMaxLevels; ++L) {
        bool found = false;
        for (int Kl = 1; Kl <= T; ++Kl)
            for (int n = L == 1 ? 1 : 2;; ++n) {
                long K0 = Kl;
                for (int l = 1; l < L; ++l) K0 *= n;
                if (L > 1 && K0 >= 2L * T) break;
                const int n0 = (int)((T + K0 - 1) / K0);
                const size_t b = (size_t)n0 * S + (size_t)(L - 1) * (n - 1) * S + (size_t)Kl * H;
                if (b <= B && (!found || b < best.bytes)) {
                    best = CkptPlan{};
                    best.L = L;
                    best.n = n;
                    best.n0 = n0;
                    best.bytes = b;
                    long k = K0;
                    for (int l = 0; l < L; ++l, k /= n) best.K[l] = (int)k;
                    found = true;
                }
                if (L == 1) break;
            }
        if (found) return true;
    }
    return false;
}

static size_t state_bytes(const std::vector<std::pair<float *, size_t>> &st) {
    size_t b = 0;
    for (auto &e : st) b += e.second;
    return b;
}
static int state_save(const std::vector<std::pair<float *, size_t>> &st, char *dst, hipStream_t s) {
    for (auto &e : st) {
        HIP_TRY(hipMemcpyAsync(dst, e.first, e.second, hipMemcpyDeviceToDevice, s));
        dst += e.second;
    }
    return 0;
}
static int state_load(const std::vector<std::pair<float *, size_t>> &st, const char *src,
               hipStream_t s) {
    for (auto &e : st) {
        HIP_TRY(hipMemcpyAsync(e.first, src, e.second, hipMemcpyDeviceToDevice, s));
        src += e.second;
    }
    return 0;
}

// One forward step without receivers (the recomputations).
static void step_forward(seis_model *m, int t, hipStream_t st) {
    step_sources(m, t, 1, st);
    step_v(m, st);
    step_s(m, st);
    if (m->c.freesurf && !m->a.fuse_surface) step_surface(m, st);
}

// Set-up of a viscoelastic gradient run: the moduli of the gradient
// formulas, the checkpoint plan and its buffers.
static int visco_prepare(seis_model *m) {
    auto &G = m->grad;
    const seis_csts &c = m->c;
    const int tmax = tmax_of(m);
    hipStream_t st = m->ctx->stream;
    // moduli of the gradient formulas: at f0 (FIX_VISCO_F0) or relaxed
    const bool f0 = c.gradfix & FIX_VISCO_F0;
    const std::vector<float> &Ms = f0 ? m->rp.Mf0 : m->rp.M, &mus = f0 ? m->rp.muf0 : m->rp.mu;
    std::vector<float> h;
    std::vector<double> hd;
    for (int q = 0; q < 2; ++q) {
        to_padded(m->g, (q ? mus : Ms).data(), 1.0f, h);
        const void *src = h.data();
        size_t bytes = h.size() * 4;
        if (m->prec == P_FP64) {   // FP64 runs read the moduli in double
            hd.assign(h.begin(), h.end());
            src = hd.data();
            bytes = hd.size() * 8;
        }
        HIP_TRY(hipMemcpyAsync(q ? G.mus : G.Ms, src, bytes, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    const auto state = forward_state(m);
    const size_t sbytes = state_bytes(state);
    const size_t hbytes = (size_t)G.nvb * m->g.nfield * m->selem;
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    const size_t held = G.ckpt_bytes + G.hist_bytes;
    const size_t budget = G.mem_limit ? G.mem_limit : held + freeb / 10 * 9;
    CkptPlan P;
    if (!plan_checkpoints(tmax, sbytes, hbytes, budget, P)) {
        char buf[240];
        std::snprintf(buf, sizeof(buf),
                      "viscoelastic gradient: no checkpoint plan for %d steps fits %.2f GB "
                      "(state %.2f GB, history step %.2f GB)", tmax, budget / 1e9, sbytes / 1e9,
                      hbytes / 1e9);
        return fail(buf);
    }
    const size_t ckb = ((size_t)P.n0 + (size_t)(P.L - 1) * (P.n - 1)) * sbytes;
    const size_t hib = (size_t)P.K[P.L - 1] * hbytes;
    if (G.ckpt_bytes + G.hist_bytes > budget) {   // a larger earlier plan: give it back
        if (grad_buffer(m, &G.ckpt, &G.ckpt_bytes, 0, true)) return 1;
        if (grad_buffer(m, &G.hist, &G.hist_bytes, 0, true)) return 1;
    }
    if (grad_buffer(m, &G.ckpt, &G.ckpt_bytes, ckb, false)) return 1;
    if (grad_buffer(m, &G.hist, &G.hist_bytes, hib, false)) return 1;
    G.plan = P;
    return 0;
}

// Viscoelastic forward run for the gradient: the level-1 checkpoints of the
// plan, then the seismograms.
static int forward_save_visco(seis_model *m, float *out) {
    auto &G = m->grad;
    const seis_csts &c = m->c;
    const int tmax = tmax_of(m);
    hipStream_t st = m->ctx->stream;
    if (visco_prepare(m)) return 1;
    const CkptPlan &P = G.plan;
    const auto state = forward_state(m);
    const size_t sbytes = state_bytes(state);
    for (int t = 0; t < tmax; ++t) {
        if (t % P.K[0] == 0 &&
            state_save(state, (char *)G.ckpt + (size_t)(t / P.K[0]) * sbytes, st))
            return 1;
        step_sources(m, t, 1, st);
        step_v(m, st);
        step_s(m, st);
        if (c.freesurf && !m->a.fuse_surface) step_surface(m, st);
        step_varsout(m, t, st);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    G.saved = true;
    return out ? seis_shot_read(m, out) : 0;
}

// Movie of a gradient run (movout with gradout, src/time_stepping.c:926-929):
// after the adjoint step t with t % movout == 0 the forward wavefields go to
// frame (t+1)/movout - 1 (the frame index of src/time_stepping.c:317-320),
// interior cells only, frames [nvar][NT/movout][NX][NY][NZ] (z fastest).
struct MovieOut {
    int movout = 0, nvar = 0;
    const char *const *vars = nullptr;
    float *frames = nullptr;
};

static int movie_snap(seis_model *m, const MovieOut *mv, int t) {
    if (!mv || !mv->frames || t % mv->movout) return 0;
    const int nfr = m->c.NT / mv->movout, f = (t + 1) / mv->movout - 1;
    if (f < 0 || f >= nfr) return 0;
    const Layout &g = m->g;
    const int H = g.H, Hy = g.nd == 3 ? H : 0;
    const long NZp = g.NZ + 2 * H, NYp = g.NY + 2 * Hy, NXp = g.NX + 2 * H;
    const size_t ncell = (size_t)g.NZ * g.NY * g.NX;
    std::vector<float> pad((size_t)NZp * NYp * NXp);
    for (int v = 0; v < mv->nvar; ++v) {
        if (seis_read_field(m, mv->vars[v], pad.data())) return 1;
        float *dst = mv->frames + ((size_t)v * nfr + f) * ncell;
        for (long x = 0; x < g.NX; ++x)
            for (long y = 0; y < g.NY; ++y)
                for (long z = 0; z < g.NZ; ++z)
                    dst[((size_t)x * g.NY + y) * g.NZ + z] =
                        pad[((x + H) * NYp + y + Hy) * NZp + z + H];
    }
    return 0;
}

// Viscoelastic adjoint over the plan of forward_save_visco (R: float, or
// double in FP64 runs).
extern "C++" {
template <typename R>
struct ViscoAdj {
    seis_model *m;
    AdjArgs<R> *A;
    std::vector<std::pair<float *, size_t>> state;
    std::vector<float *> hf;    // forward fields the adjoint step reads
    size_t sbytes, fb;          // bytes of a checkpoint, of one stored field
    int tmin, serial;
    bool half, prof;
    size_t k = 0;               // adjoint steps done (event slots)
    const MovieOut *mv = nullptr;

    // level-l segment [t0, t1), whose starting state is current
    int segment(int l, int t0, int t1) {
        auto &G = m->grad;
        const CkptPlan &P = G.plan;
        hipStream_t st = m->ctx->stream;
        if (l == P.L - 1) return innermost(t0, t1);
        const int K = P.K[l + 1], nsub = (t1 - t0 + K - 1) / K;
        char *slots = (char *)G.ckpt + ((size_t)P.n0 + (size_t)l * (P.n - 1)) * sbytes;
        for (int t = t0; t < t0 + (nsub - 1) * K; ++t) {
            if ((t - t0) % K == 0 && state_save(state, slots + (size_t)((t - t0) / K) * sbytes, st))
                return 1;
            step_forward(m, t, st);
        }
        for (int s = nsub - 1; s >= 0; --s) {
            const int u0 = t0 + s * K, u1 = std::min(u0 + K, t1);
            if (u1 - 1 <= tmin) break;
            if (s < nsub - 1 && state_load(state, slots + (size_t)s * sbytes, st)) return 1;
            if (segment(l + 1, u0, u1)) return 1;
        }
        return 0;
    }

    // recompute [t0, t1) keeping every post-source state, then the adjoint
    // steps t1-1 .. max(t0, tmin+1) over them
    int innermost(int t0, int t1) {
        auto &G = m->grad;
        const seis_csts &c = m->c;
        hipStream_t st = m->ctx->stream;
        const int NT = c.NT, nout = (int)m->outn.size();
        for (int t = t0; t < t1; ++t) {
            // movie frame: the recomputed forward state at the start of step
            // t, what backpropagation reconstructs after the adjoint step t
            if (t > tmin && movie_snap(m, mv, t)) return 1;
            step_sources(m, t, 1, st);
            char *dst = (char *)G.hist + (size_t)(t - t0) * G.nvb * fb;
            for (size_t q = 0; q < hf.size(); ++q)
                HIP_TRY(hipMemcpyAsync(dst + q * fb, hf[q], fb, hipMemcpyDeviceToDevice, st));
            step_v(m, st);
            step_s(m, st);
            if (c.freesurf && !m->a.fuse_surface) step_surface(m, st);
        }
        for (int t = t1 - 1; t >= std::max(t0, tmin + 1); --t) {
            if (c.freesurf)
                launch_surface_adj(A->r, (R *)G.stmp, m->nd_case, c.gradfix & FIX_SURF_SYY, st);
            launch_residuals(A->r, m->nd_case, m->shot, m->out_codes.data(), nout, G.res, NT, t,
                             serial, st);
            const char *h = (const char *)G.hist + (size_t)(t - t0) * G.nvb * fb;
            for (size_t q = 0; q < hf.size(); ++q) A->st[q] = h + q * fb;
            if (prof) HIP_TRY(hipEventRecord(m->evpool[2 + 4 * k], st));
            launch_adjh_s(*A, m->nd_case, c.L, half, st);
            if (prof) HIP_TRY(hipEventRecord(m->evpool[3 + 4 * k], st));
            if (prof) HIP_TRY(hipEventRecord(m->evpool[4 + 4 * k], st));
            launch_adjh_v(*A, m->nd_case, half, st);
            if (prof) HIP_TRY(hipEventRecord(m->evpool[5 + 4 * k], st));
            ++k;
        }
        return 0;
    }
};

template <typename R>
static int adjoint_visco(seis_model *m, AdjArgs<R> &A, int serial, const MovieOut *mv) {
    auto &G = m->grad;
    const CkptPlan &P = G.plan;
    const int tmax = tmax_of(m), tmin = m->c.tmin;
    hipStream_t st = m->ctx->stream;
    ViscoAdj<R> V;
    V.m = m;
    V.A = &A;
    V.state = forward_state(m);
    V.hf = history_fields(m);
    V.sbytes = state_bytes(V.state);
    V.fb = (size_t)m->g.nfield * m->selem;
    V.tmin = tmin;
    V.serial = serial;
    V.half = m->prec == P_FP16;
    V.prof = m->profiling;
    V.mv = mv;
    A.Ms = (const R *)G.Ms;
    A.mus = (const R *)G.mus;
    // with profiling on, events around the two adjoint kernels of every step
    const size_t nst = (size_t)std::max(0, tmax - 1 - tmin);
    hipEvent_t e0 = pool_event(m, 0), e1 = pool_event(m, 1);
    if (!e0 || !e1 || (V.prof && !pool_event(m, 1 + 4 * nst)))
        return fail("hipEventCreate failed");
    HIP_TRY(hipEventRecord(e0, st));
    for (int s0 = P.n0 - 1; s0 >= 0; --s0) {
        const int t0 = s0 * P.K[0], t1 = std::min(t0 + P.K[0], tmax);
        if (t1 - 1 <= tmin) break;
        if (state_load(V.state, (const char *)G.ckpt + (size_t)s0 * V.sbytes, st)) return 1;
        if (V.segment(0, t0, t1)) return 1;
    }
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    m->ms_loop = ms;
    m->ms_v = m->ms_s = 0;
    if (V.prof)
        for (size_t q = 0; q < V.k; ++q) {
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, m->evpool[2 + 4 * q], m->evpool[3 + 4 * q]));
            HIP_TRY(hipEventElapsedTime(&b, m->evpool[4 + 4 * q], m->evpool[5 + 4 * q]));
            m->ms_s += a;
            m->ms_v += b;
        }
    return 0;
}
}  // extern "C++"

}  // namespace

}  // extern "C"

namespace {
// Boundary-shell buffer of an elastic gradient run, tmax steps of every
// saved wavefield.
static int bnd_prepare(seis_model *m) {
    auto &G = m->grad;
    const long need = (long)tmax_of(m);
    if (G.bnd_steps >= need) return 0;
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    const size_t be = m->prec == P_FP64 ? 8 : 4;   // bytes of a saved value
    if (G.bnd) {
        m->bytes -= (size_t)G.bnd_steps * G.nvb * G.sh.n * G.bnd_elem;
        m->ctx->release(G.bnd);
        G.bnd = nullptr;
        G.bnd_steps = 0;
    }
    void *p = nullptr;
    const size_t bytes = (size_t)need * G.nvb * G.sh.n * be;
    if (m->ctx->alloc(bytes, &p)) return 1;
    G.bnd = (float *)p;
    G.bnd_steps = need;
    G.bnd_elem = be;
    m->bytes += bytes;
    return 0;
}

// Header of a saved forward run copied to the host (seis_grad_saved_read).
struct SavedHeader {
    char magic[8];          // "SEISCKP1"
    int32_t kind;           // 0: elastic (final state + boundary), 1: viscoelastic
    int32_t tmax, n0, K0;
    uint64_t state_bytes, payload_bytes;
};
static_assert(sizeof(SavedHeader) % 4 == 0, "float-sized header");

static int saved_layout(seis_model *m, SavedHeader &h) {
    auto &G = m->grad;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "SEISCKP1", 8);
    const auto state = forward_state(m);
    h.state_bytes = state_bytes(state);
    h.tmax = tmax_of(m);
    if (G.visco) {
        h.kind = 1;
        h.n0 = G.plan.n0;
        h.K0 = G.plan.K[0];
        h.payload_bytes = (uint64_t)h.n0 * h.state_bytes;
    } else {
        h.kind = 0;
        h.payload_bytes = h.state_bytes + (uint64_t)h.tmax * G.nvb * G.sh.n *
                                              (m->prec == P_FP64 ? 8 : 4);
    }
    return 0;
}

// ---- DFT-domain gradient (back_prop_type 2) --------------------------------
// Sampling of src/assign_modeling_case.c:549-570: every DTNYQ steps from
// tmin, NTNYQ samples, frequency bins of the record at df.
int dft_sampling(seis_model *m) {
    auto &D = m->dft;
    const seis_csts &c = m->c;
    float fmax = 0;
    for (float f : D.freqs) fmax = std::max(fmax, f);
    const int tmax = tmax_of(m);
    D.dtnyq = (int)std::ceil(0.0156 / fmax / c.dt);
    D.ntnyq = (tmax - c.tmin) / D.dtnyq + 1;
    D.df = (float)(1.0 / D.ntnyq / c.dt / D.dtnyq);
    D.bins.resize(D.nf);
    D.w.resize(D.nf);
    for (int f = 0; f < D.nf; ++f) {
        D.bins[f] = (float)std::floor(D.freqs[f] / D.df);
        D.w[f] = (float)(2.0 * kPiDft * D.df * D.bins[f]);
    }
    return 0;
}

// The arrays whose spectra are kept, in the order of the gradient formula.
DftArrays dft_arrays(const seis_model *m, const FDArgs<float> &a) {
    DftArrays q;
    const long nf = m->g.nfield;
    const int L = m->c.L;
    if (m->c.ND == 3) {
        const float *f[9] = {a.vx, a.vy, a.vz, a.sxx, a.syy, a.szz, a.sxy, a.syz, a.sxz};
        for (const float *p : f) q.p[q.n++] = p;
        const float *r[6] = {a.rxx, a.ryy, a.rzz, a.rxy, a.ryz, a.rxz};
        for (int k = 0; k < 6 && L > 0; ++k)
            for (int l = 0; l < L; ++l) q.p[q.n++] = r[k] + (long)l * nf;
    } else {
        const float *f[5] = {a.vx, a.vz, a.sxx, a.szz, a.sxz};
        for (const float *p : f) q.p[q.n++] = p;
        const float *r[3] = {a.rxx, a.rzz, a.rxz};
        for (int k = 0; k < 3 && L > 0; ++k)
            for (int l = 0; l < L; ++l) q.p[q.n++] = r[k] + (long)l * nf;
    }
    return q;
}

// Spectra, factor table and cell parameters of the model (once per
// frequency set); the spectra are zeroed for each forward run.
int dft_prepare(seis_model *m) {
    auto &D = m->dft;
    auto &G = m->grad;
    const seis_csts &c = m->c;
    dft_sampling(m);
    D.nq = c.ND == 3 ? 9 + 6 * c.L : 5 + 3 * c.L;
    hipStream_t st = m->ctx->stream;
    const long nfd = m->g.nfield;
    if (D.cap != (size_t)D.nf || !D.Ff) {
        // buffers of an earlier frequency set stay with the model until it
        // is destroyed
        const long n = 2L * D.nq * D.nf * nfd;
        if (!(D.Ff = m->dalloc(n)) || !(D.Fa = m->dalloc(n))) return 1;
        if (!(D.w_dev = m->dalloc(D.nf))) return 1;
        if (!(D.fact = m->dalloc(2L * D.ntnyq * D.nf))) return 1;
        D.cap = D.nf;
    }
    std::vector<float> fact(2 * (size_t)D.ntnyq * D.nf);
    const float sc = (float)D.dtnyq * c.dt;
    for (int th = 0; th < D.ntnyq; ++th)
        for (int f = 0; f < D.nf; ++f) {
            const double a = 2.0 * D.bins[f] * th / D.ntnyq;
            fact[2 * ((size_t)th * D.nf + f)] = (float)(sc * std::cos(kPiDft * a));
            fact[2 * ((size_t)th * D.nf + f) + 1] = (float)(-sc * std::sin(kPiDft * a));
        }
    HIP_TRY(hipMemcpyAsync(D.fact, fact.data(), fact.size() * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(D.w_dev, D.w.data(), D.nf * 4, hipMemcpyHostToDevice, st));
    const size_t n = (size_t)m->g.NZ * m->g.NY * m->g.NX;
    if (!D.par && !(D.par = m->dalloc(5L * n))) return 1;
    std::vector<float> par(5 * n, 0.f);
    const RefParams &p = m->rp;
    std::copy(p.M.begin(), p.M.end(), par.begin());
    if (!p.mu.empty()) std::copy(p.mu.begin(), p.mu.end(), par.begin() + n);
    if (c.L > 0) {
        std::copy(p.taup.begin(), p.taup.end(), par.begin() + 2 * n);
        std::copy(p.taus.begin(), p.taus.end(), par.begin() + 3 * n);
    }
    std::copy(p.rho.begin(), p.rho.end(), par.begin() + 4 * n);
    HIP_TRY(hipMemcpyAsync(D.par, par.data(), par.size() * 4, hipMemcpyHostToDevice, st));
    launch_fill(D.Ff, 8 * (size_t)D.nq * D.nf * nfd, st);
    launch_fill(D.Fa, 8 * (size_t)D.nq * D.nf * nfd, st);
    // tau gradients of every viscoelastic parameterisation
    if (c.L > 0 && !G.gtp) {
        const long nd2 = 2L * nfd;
        for (double **q : {&G.gtp, &G.gts}) {
            if (!(*q = (double *)m->dalloc(nd2))) return 1;
            launch_fill(*q, 8 * (size_t)nfd, st);
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// Forward run with the spectra of the state before the sources of every
// sampled step (src/time_stepping.c:767-780).
int forward_dft(seis_model *m, float *out) {
    auto &D = m->dft;
    const seis_csts &c = m->c;
    const int tmax = tmax_of(m), tmin = c.tmin;
    hipStream_t st = m->ctx->stream;
    const DftArrays q = dft_arrays(m, m->a);
    for (int t = 0; t < tmax; ++t) {
        if (t >= tmin && (t - tmin) % D.dtnyq == 0)
            launch_dft_acc(q, D.Ff, D.fact + 2L * ((t - tmin) / D.dtnyq) * D.nf, D.nf, m->g, st);
        step_sources(m, t, 1, st);
        step_v(m, st);
        step_s(m, st);
        if (c.freesurf && !m->a.fuse_surface) step_surface(m, st);
        step_varsout(m, t, st);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    m->grad.saved = true;
    return out ? seis_shot_read(m, out) : 0;
}

// Adjoint run of the adjoint wavefield alone (src/time_stepping.c:869-925
// with BACK_PROP_TYPE 2: surface_adj, residuals, the adjoint updates, the
// spectra after them), then the gradient from the two sets of spectra.
int adjoint_dft(seis_model *m) {
    auto &D = m->dft;
    auto &G = m->grad;
    const seis_csts &c = m->c;
    const int NT = c.NT, nout = (int)m->outn.size(), tmax = tmax_of(m), tmin = c.tmin;
    hipStream_t st = m->ctx->stream;
    FDArgs<float> r = adjoint_args(m);
    r.fuse_surface = 0;
    const DftArrays q = dft_arrays(m, r);
    const int serial = receivers_share_cells(m, m->rcell_host);
    for (int t = tmax - 1; t > tmin; --t) {
        if (c.freesurf) launch_surface_adj(r, G.stmp, m->nd_case, c.gradfix & FIX_SURF_SYY, st);
        launch_residuals(r, m->nd_case, m->shot, m->out_codes.data(), nout, G.res, NT, t, serial,
                         st);
        launch_update_s(r, m->nd_case, st);
        launch_update_v(r, m->nd_case, st);
        if ((t - tmin) % D.dtnyq == 0)
            launch_dft_acc(q, D.Fa, D.fact + 2L * ((t - tmin) / D.dtnyq) * D.nf, D.nf, m->g, st);
    }
    DftGrad d;
    d.Ff = D.Ff;
    d.Fa = D.Fa;
    d.par = D.par;
    d.w = D.w_dev;
    d.gM = G.gM;
    d.gmu = G.gmu;
    d.grho = G.grho;
    d.gtp = G.gtp;
    d.gts = G.gts;
    d.nf = D.nf;
    d.L = c.L;
    d.pt = c.param_type;
    d.ntnyq = D.ntnyq;
    float al = 0;
    const float w0 = 2.0 * kPiDft * c.f0;
    for (int l = 0; l < c.L; ++l) {
        d.tausig[l] = 1.0 / (2.0 * kPiDft * c.FL[l]);
        const double rr = w0 / (2.0 * kPiDft * c.FL[l]);
        al += rr * rr / (1.0 + rr * rr);
    }
    d.al = al;
    launch_dft_grad(d, m->g, st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    G.saved = false;
    return 0;
}
}  // namespace

extern "C" {

int seis_grad_dft(seis_model *m, int nfreqs, const float *freqs) {
    g_err.clear();
    if (!m) return fail("seis_grad_dft: null model");
    if (nfreqs == 0) {
        m->dft.nf = 0;
        return 0;
    }
    const seis_csts &c = m->c;
    if (nfreqs < 0 || nfreqs > kMaxDftFreqs || !freqs)
        return fail("seis_grad_dft: give 1 to 512 gradient frequencies");
    for (int f = 0; f < nfreqs; ++f)
        if (!(freqs[f] > 0)) return fail("seis_grad_dft: gradient frequencies must be positive");
    if (c.ND != 2 && c.ND != 3)
        return fail("seis_grad_dft: the DFT gradient needs the elastic or viscoelastic case "
                    "(ND 2 or 3)");
    if (m->prec != P_FP32) return fail("seis_grad_dft: the DFT gradient runs in FP32 (FP16=0)");
    if (c.param_type == 3 && c.L == 0)
        return fail("seis_grad_dft: param_type 3 needs a viscoelastic model (L > 0)");
    if (c.restype != 0 || c.Hout || c.scalerms || c.scaleshot || c.scalermsnorm)
        return fail("seis_grad_dft: the DFT gradient takes the L2 misfit without trace scaling, "
                    "no Hessian and no RTM (restype 0, Hout 0, scalerms/scaleshot/scalermsnorm 0)");
    m->dft.freqs.assign(freqs, freqs + nfreqs);
    m->dft.nf = nfreqs;
    dft_sampling(m);
    // bin 0 (a frequency below the record's step df) would divide the
    // memory-variable terms by zero
    for (float b : m->dft.bins)
        if (b < 1) {
            m->dft.nf = 0;
            return fail("seis_grad_dft: gradient frequencies below the frequency step of the "
                        "record (bin 0)");
        }
    // the rms bins must stay below the Nyquist frequency of the record
    const int tmax = tmax_of(m), nfft = fft_fast_size(tmax);
    for (float f : m->dft.freqs)
        if ((int)(f * nfft * c.dt + 1) > nfft / 2) {
            m->dft.nf = 0;
            return fail("seis_grad_dft: gradient frequencies above the Nyquist frequency of the "
                        "record");
        }
    return 0;
}

int seis_forward_save(seis_model *m, int nsrc, const float *src_pos, const float *src, int nrec,
                      const float *rec_pos, float *out) {
    g_err.clear();
    if (grad_setup(m)) return 1;
    auto &G = m->grad;
    if (m->dft.nf) {
        if (dft_prepare(m)) return 1;
        if (seis_shot_bind(m, nsrc, src_pos, src, nrec, rec_pos)) return 1;
        if (seis_launch_init(m)) return 1;
        return forward_dft(m, out);
    }
    if (G.visco) {
        if (seis_shot_bind(m, nsrc, src_pos, src, nrec, rec_pos)) return 1;
        if (seis_launch_init(m)) return 1;
        return forward_save_visco(m, out);
    }
    const int tmax = tmax_of(m);
    if (bnd_prepare(m)) return 1;
    if (seis_shot_bind(m, nsrc, src_pos, src, nrec, rec_pos)) return 1;
    if (seis_launch_init(m)) return 1;
    hipStream_t st = m->ctx->stream;
    const size_t per = (size_t)G.nvb * G.sh.n;
    // forward step order with boundary saving after update_s
    // (src/time_stepping.c:785-816)
    // the boundary is saved before the free surface is applied, so the
    // surface runs as its own kernel here
    const int fused = m->a.fuse_surface;
    m->a.fuse_surface = 0;
    m->ad.fuse_surface = 0;   // the FP64 view of the same arguments
    // step t (relative to *tb when tb is given): varsout of t-1 with the
    // sources of t, the updates, the shell of t, the free surface
    auto step = [&](int t, const int *tb) {
        step_io(m, t - 1, t, st, tb);
        step_v(m, st);
        step_s(m, st);
        if (m->prec == P_FP64)
            launch_savebnd(m->ad, G.sh, m->nd_case, (double *)G.bnd + (long)t * (long)per, st,
                           tb, (long)per);
        else
            launch_savebnd(m->a, G.sh, m->nd_case, G.bnd + (long)t * (long)per, st, tb,
                           (long)per);
        if (m->c.freesurf) step_surface(m, st);
    };
    int t0 = 0;
    // 2D: runs of kGraphSteps steps replayed from a captured graph
    if (m->c.ND != 3 && !m->profiling && tmax >= kGraphSteps && graph_enabled()) {
        if (step_counter(m)) return 1;
        uint64_t key = step_key(m);
        key = key_mix(key, &G.bnd, sizeof(G.bnd));
        key = key_mix(key, &G.sh, sizeof(G.sh));
        key = key_mix(key, &per, sizeof(per));
        if (capture_steps(m->fs_graph, key, st, [&] {
                for (int k = 0; k < kGraphSteps; ++k) step(k, m->sg_t);
                launch_tcounter(m->sg_t, kGraphSteps, 1, st);
            }))
            return 1;
        launch_tcounter(m->sg_t, 0, 0, st);
        for (; t0 + kGraphSteps <= tmax; t0 += kGraphSteps)
            HIP_TRY(hipGraphLaunch(m->fs_graph.e, st));
    }
    for (int t = t0; t < tmax; ++t) step(t, nullptr);
    step_io(m, tmax - 1, -1, st);   // varsout of the last step
    m->a.fuse_surface = fused;
    m->ad.fuse_surface = fused;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    G.saved = true;
    return out ? seis_shot_read(m, out) : 0;
}

int seis_grad_saved_size(seis_model *m, size_t *bytes) {
    g_err.clear();
    if (!m->grad.saved) return fail("seis_grad_saved_size: no saved forward run is bound");
    SavedHeader h;
    saved_layout(m, h);
    *bytes = sizeof(h) + h.payload_bytes;
    return 0;
}

int seis_grad_saved_read(seis_model *m, float *host) {
    g_err.clear();
    auto &G = m->grad;
    if (!G.saved) return fail("seis_grad_saved_read: no saved forward run is bound");
    SavedHeader h;
    saved_layout(m, h);
    std::memcpy(host, &h, sizeof(h));
    char *dst = (char *)host + sizeof(h);
    hipStream_t st = m->ctx->stream;
    if (h.kind == 1) {
        if (h.payload_bytes)
            HIP_TRY(hipMemcpyAsync(dst, G.ckpt, h.payload_bytes, hipMemcpyDeviceToHost, st));
    } else {
        for (auto &e : forward_state(m)) {
            HIP_TRY(hipMemcpyAsync(dst, e.first, e.second, hipMemcpyDeviceToHost, st));
            dst += e.second;
        }
        const size_t nb = h.payload_bytes - h.state_bytes;
        if (nb) HIP_TRY(hipMemcpyAsync(dst, G.bnd, nb, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int seis_grad_saved_write(seis_model *m, const float *host) {
    g_err.clear();
    if (!m->shot_bound) return fail("seis_grad_saved_write: bind the shot first");
    if (grad_setup(m)) return 1;
    auto &G = m->grad;
    G.saved = false;
    SavedHeader h;
    std::memcpy(&h, host, sizeof(h));
    if (std::memcmp(h.magic, "SEISCKP1", 8))
        return fail("seis_grad_saved_write: not a saved forward run of this library");
    if (G.visco && visco_prepare(m)) return 1;
    if (!G.visco && bnd_prepare(m)) return 1;
    SavedHeader want;
    saved_layout(m, want);
    if (h.kind != want.kind || h.tmax != want.tmax || h.state_bytes != want.state_bytes ||
        h.n0 != want.n0 || h.K0 != want.K0 || h.payload_bytes != want.payload_bytes)
        return fail("seis_grad_saved_write: the saved forward run does not match this model "
                    "(kind, time steps, state size or checkpoint plan differ)");
    const char *src = (const char *)host + sizeof(h);
    hipStream_t st = m->ctx->stream;
    if (h.kind == 1) {
        if (h.payload_bytes)
            HIP_TRY(hipMemcpyAsync(G.ckpt, src, h.payload_bytes, hipMemcpyHostToDevice, st));
    } else {
        for (auto &e : forward_state(m)) {
            HIP_TRY(hipMemcpyAsync(e.first, src, e.second, hipMemcpyHostToDevice, st));
            src += e.second;
        }
        const size_t nb = h.payload_bytes - h.state_bytes;
        if (nb) HIP_TRY(hipMemcpyAsync(G.bnd, src, nb, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    G.saved = true;
    return 0;
}

int seis_misfit(const seis_csts *c, int nout, int nrec, const float *out, const float *din,
                float *res, double *rms, double *rmsnorm) {
    g_err.clear();
    if (!c || nout < 0 || nrec < 0 || c->NT <= 0) return fail("seis_misfit: bad sizes");
    const double r = residuals(*c, nout, nrec, out, din, res, rmsnorm);
    if (rms) *rms = r;
    return 0;
}

int seis_residuals(const seis_model *m, int nrec, const float *out, const float *din, float *res,
                   double *rms, double *rmsnorm) {
    double r = residuals(m->c, (int)m->outn.size(), nrec, out, din, res, rmsnorm);
    if (m->dft.nf) r = dft_rms(m->c, (int)m->outn.size(), nrec, out, din, m->dft.freqs, rmsnorm);
    if (rms) *rms = r;
    return 0;
}

int seis_res_scale(const seis_model *m, int nrec, const float *rec_pos, float *res) {
    res_scale(m->c, m->rp, m->outn, nrec, rec_pos, res, m->c.gradfix & FIX_RES_SCALE);
    return 0;
}

namespace {
// The adjoint steps of adjoint_run once the adjoint sources are on the
// device: R = float, or double in FP64 runs (FP16=64).
extern "C++" {
template <typename R>
static int adjoint_loops(seis_model *m, const MovieOut *mv) {
    auto &G = m->grad;
    const int NT = m->c.NT, nout = (int)m->outn.size();
    const int tmax = tmax_of(m), tmin = m->c.tmin;
    hipStream_t st = m->ctx->stream;
    const bool scaled = m->prec == P_SCALED || m->prec == P_FP16;
    AdjArgs<R> A;
    if constexpr (sizeof(R) == 8) {
        A.f = m->ad;
        A.r = as_double(adjoint_args(m));
    } else {
        A.f = m->a;
        A.r = adjoint_args(m);
    }
    A.gM = G.gM;
    A.gmu = G.gmu;
    A.grho = G.grho;
    A.sh = G.sh;
    if (scaled) {
        // forward stresses are stored times as, adjoint ones as is;
        // forward velocities times av, adjoint ones times q
        A.gscale_s = 1.0 / m->as;
        A.gscale_v = 1.0 / ((double)m->av * m->q);
        A.hscale = 1.0 / ((double)m->as * m->as);
    }
    A.rtm = m->c.restype == 1;
    A.hM = G.hM;
    A.hmu = G.hmu;
    if (G.gtp && !A.rtm) {
        A.gtp = G.gtp;
        A.gts = G.gts;
        A.s0 = (R)relax_sum(m->c);
    }
    if (G.grz && !A.rtm) {
        for (int q = 0; q < 3; ++q) A.gsh[q] = G.gsh[q];
        A.gry = G.gry;
        A.grz = G.grz;
    }
    A.hfix = (m->c.gradfix & FIX_HESS) != 0;
    if (m->c.freesurf && (!A.r.M || !A.r.mu || !G.stmp))
        return fail("seis_adjoint: the adjoint free surface has no moduli on the device");
    const int serial = receivers_share_cells(m, m->rcell_host);
    if (G.visco) {
        const int r = adjoint_visco(m, A, serial, mv);
        G.saved = false;
        return r;
    }
    const size_t per = (size_t)G.nvb * G.sh.n;
    // with profiling on, events around the two fused kernels of every step
    const bool prof = m->profiling;
    const size_t nst = (size_t)(tmax - 1 - tmin);
    hipEvent_t e0 = pool_event(m, 0), e1 = pool_event(m, 1);
    if (!e0 || !e1 || (prof && !pool_event(m, 1 + 4 * nst))) return fail("hipEventCreate failed");
    HIP_TRY(hipEventRecord(e0, st));
    size_t k = 0;
    const R *bnd0 = (const R *)G.bnd;
    // adjoint step t (relative to *tb when tb is given; the shell of step
    // t-1), src/time_stepping.c:869-940
    auto step = [&](int t, const int *tb) {
        if (m->c.freesurf)
            launch_surface_adj(A.r, (R *)G.stmp, m->nd_case, m->c.gradfix & FIX_SURF_SYY, st);
        launch_residuals(A.r, m->nd_case, m->shot, m->out_codes.data(), nout, G.res, NT, t, serial,
                         st, tb);
        A.tb = tb;
        A.bper = (long)per;
        A.bofs = t - 1;
        A.bnd = tb ? bnd0 : bnd0 + (long)(t - 1) * (long)per;
        launch_adj_s(A, m->nd_case, st);
        launch_adj_v(A, m->nd_case, st);
        step_sources(m, t, -1, st, tb);
    };
    int t = tmax - 1;
    // 2D without profiling or movie: runs of kGraphSteps adjoint steps
    // replayed from a captured graph, the counter holding the run's first t
    if (m->c.ND != 3 && !prof && !mv && tmax - 1 - tmin >= kGraphSteps && graph_enabled()) {
        if (step_counter(m)) return 1;
        AdjArgs<R> K = A;
        K.bnd = bnd0;
        K.tb = m->sg_t;
        K.bper = (long)per;
        uint64_t key = step_key(m);
        key = key_mix(key, &K, sizeof(K));
        key = key_mix(key, &G.res, sizeof(G.res));
        key = key_mix(key, &G.stmp, sizeof(G.stmp));
        key = key_mix(key, &serial, sizeof(serial));
        key = key_mix(key, &NT, sizeof(NT));
        const int ss = sizeof(R);
        key = key_mix(key, &ss, sizeof(ss));
        if (capture_steps(m->ad_graph, key, st, [&] {
                for (int j = 0; j < kGraphSteps; ++j) step(-j, m->sg_t);
                launch_tcounter(m->sg_t, -kGraphSteps, 1, st);
            }))
            return 1;
        launch_tcounter(m->sg_t, t, 0, st);
        for (; t - kGraphSteps >= tmin; t -= kGraphSteps, k += kGraphSteps)
            HIP_TRY(hipGraphLaunch(m->ad_graph.e, st));
    }
    for (; t > tmin; --t, ++k) {
        if (!prof) {
            step(t, nullptr);
        } else {
            if (m->c.freesurf)
                launch_surface_adj(A.r, (R *)G.stmp, m->nd_case, m->c.gradfix & FIX_SURF_SYY, st);
            launch_residuals(A.r, m->nd_case, m->shot, m->out_codes.data(), nout, G.res, NT, t,
                             serial, st);
            A.tb = nullptr;
            A.bnd = bnd0 + (long)(t - 1) * (long)per;
            HIP_TRY(hipEventRecord(m->evpool[2 + 4 * k], st));
            launch_adj_s(A, m->nd_case, st);
            HIP_TRY(hipEventRecord(m->evpool[3 + 4 * k], st));
            HIP_TRY(hipEventRecord(m->evpool[4 + 4 * k], st));
            launch_adj_v(A, m->nd_case, st);
            HIP_TRY(hipEventRecord(m->evpool[5 + 4 * k], st));
            step_sources(m, t, -1, st);
        }
        if (movie_snap(m, mv, t)) return 1;
    }
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(e1));
    G.saved = false;   // the forward fields are now at the start of the run
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    m->ms_loop = ms;
    m->ms_v = m->ms_s = 0;
    if (prof)
        for (size_t q = 0; q < nst; ++q) {
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, m->evpool[2 + 4 * q], m->evpool[3 + 4 * q]));
            HIP_TRY(hipEventElapsedTime(&b, m->evpool[4 + 4 * q], m->evpool[5 + 4 * q]));
            m->ms_s += a;
            m->ms_v += b;
        }
    return 0;
}

}  // extern "C++"

// dev: res is a device array; dev_scale: apply res_scale to it on the device
// (the host path receives residuals already scaled by seis_res_scale).
static int adjoint_run(seis_model *m, const float *res, const MovieOut *mv, bool dev = false,
                       bool dev_scale = false) {
    auto &G = m->grad;
    if (!G.saved) return fail("seis_adjoint: no forward run with saved boundaries is bound");
    const int NT = m->c.NT, nrec = m->shot.nrec, nout = (int)m->outn.size();
    const int tmax = tmax_of(m), tmin = m->c.tmin;
    if (tmin < 0 || tmin >= tmax) return fail("tmin outside [0, tmax)");
    hipStream_t st = m->ctx->stream;
    const size_t nres = (size_t)nout * nrec * NT;
    if (nres > G.res_cap) {
        HIP_TRY(hipStreamSynchronize(st));
        m->ctx->release(G.res);
        G.res = nullptr;
        void *p = nullptr;
        if (m->ctx->alloc(nres * 4, &p)) return 1;
        G.res = (float *)p;
        G.res_cap = nres;
    }
    // scaled wavefields (FP16=1, 2): the adjoint run keeps its stresses in
    // physical units and its velocities times q, the units the scaled
    // coefficients imply, so velocity residuals are injected times q
    const bool scaled = m->prec == P_SCALED || m->prec == P_FP16;
    std::vector<float> rs;
    if (dev && nres) {
        const size_t nfac = (size_t)nout * nrec;
        const size_t bytes = nfac * 8 + (size_t)nout * 4;
        if (bytes > G.res_fac_cap) {
            HIP_TRY(hipStreamSynchronize(st));
            m->ctx->release(G.res_fac);
            G.res_fac = nullptr;
            void *p = nullptr;
            if (m->ctx->alloc(bytes, &p)) return 1;
            G.res_fac = (double *)p;
            G.res_fac_cap = bytes;
        }
        std::vector<char> host(bytes);
        double *fac = (double *)host.data();
        float *qv = (float *)(host.data() + nfac * 8);
        if (dev_scale) {
            if (m->recpos_host.size() != (size_t)8 * nrec)
                return fail("seis_adjoint_device: no receiver positions for the bound shot");
            res_scale_factors(m->c, m->rp, m->outn, nrec, m->recpos_host.data(), fac,
                              m->c.gradfix & FIX_RES_SCALE);
        }
        for (int o = 0; o < nout; ++o) qv[o] = scaled && m->out_codes[o] <= O_VZ ? m->q : 1.0f;
        HIP_TRY(hipMemcpyAsync(G.res_fac, host.data(), bytes, hipMemcpyHostToDevice, st));
        const int blocks = (int)std::min<size_t>((nres + 255) / 256, 4096);
        hipLaunchKernelGGL(k_res_device, dim3(blocks), dim3(256), 0, st, res, G.res,
                           dev_scale ? (const double *)G.res_fac : nullptr,
                           scaled ? (const float *)(G.res_fac + nfac) : nullptr, nout, nrec, NT,
                           tmax);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));   // the host staging of the factors is released
    } else if (scaled && nres) {
        rs.assign(res, res + nres);
        for (int o = 0; o < nout; ++o)
            if (m->out_codes[o] <= O_VZ)
                for (size_t k = (size_t)o * nrec * NT; k < (size_t)(o + 1) * nrec * NT; ++k)
                    rs[k] *= m->q;
        res = rs.data();
    }
    if (nres && !dev) HIP_TRY(hipMemcpyAsync(G.res, res, nres * 4, hipMemcpyHostToDevice, st));
    if (!rs.empty()) HIP_TRY(hipStreamSynchronize(st));
    grad_zero_adjoint(m);
    if (m->dft.nf) {
        if (mv) return fail("seis_adjoint_movie: no movie of a DFT-gradient run");
        return adjoint_dft(m);
    }
    return m->prec == P_FP64 ? adjoint_loops<double>(m, mv) : adjoint_loops<float>(m, mv);
}

}  // namespace

int seis_adjoint(seis_model *m, const float *res) {
    g_err.clear();
    return adjoint_run(m, res, nullptr);
}

int seis_adjoint_device(seis_model *m, const float *dres, int res_scale_on) {
    g_err.clear();
    if (!dres && m->shot.nrec > 0 && !m->outn.empty())
        return fail("seis_adjoint_device: no residuals");
    return adjoint_run(m, dres, nullptr, true, res_scale_on != 0);
}

int seis_shot_read_device(seis_model *m, float *dout) {
    g_err.clear();
    if (!m->shot.out) return fail("seis_shot_read_device: no shot bound");
    const int NT = m->c.NT, nrec = m->shot.nrec, nout = (int)m->outn.size();
    const long n = (long)NT * nrec * nout;
    if (n == 0) return 0;
    if (!dout) return fail("seis_shot_read_device: no output array");
    hipStream_t st = m->ctx->stream;
    if (m->done_ev) HIP_TRY(hipStreamWaitEvent(st, m->done_ev, 0));
    const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
    if (m->prec == P_FP64)
        hipLaunchKernelGGL(k_seis_transpose<double>, dim3(blocks), dim3(256), 0, st,
                           (const double *)m->shot.out, dout, NT, nout, nrec);
    else
        hipLaunchKernelGGL(k_seis_transpose<float>, dim3(blocks), dim3(256), 0, st,
                           (const float *)m->shot.out, dout, NT, nout, nrec);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int seis_adjoint_movie(seis_model *m, const float *res, int movout, int nvar,
                       const char *const *vars, float *frames) {
    g_err.clear();
    if (movout <= 0 || movout > m->c.NT) return fail("seis_adjoint_movie: movout outside [1, NT]");
    if (nvar <= 0 || !vars || !frames) return fail("seis_adjoint_movie: no movie variables");
    for (int v = 0; v < nvar; ++v)
        if (!vars[v] || !field_ptr(m, vars[v]))
            return fail(std::string("seis_adjoint_movie: no wavefield named ") +
                        (vars[v] ? vars[v] : "(null)"));
    const size_t nf = (size_t)nvar * (m->c.NT / movout) * m->g.NZ * m->g.NY * m->g.NX;
    std::fill(frames, frames + nf, 0.0f);
    MovieOut mv;
    mv.movout = movout;
    mv.nvar = nvar;
    mv.vars = vars;
    mv.frames = frames;
    return adjoint_run(m, res, &mv);
}

namespace {
// Device accumulators (padded layout, double) -> parameter layout on the host.
static int accum_read(seis_model *m, const double *const *src, int k, std::vector<double> *acc) {
    const Layout &g = m->g;
    const size_t nf = (size_t)g.nfield, n = (size_t)g.NZ * g.NY * g.NX;
    std::vector<double> tmp(nf);
    for (int q = 0; q < k; ++q) {
        HIP_TRY(hipMemcpyAsync(tmp.data(), src[q], nf * 8, hipMemcpyDeviceToHost, m->ctx->stream));
        HIP_TRY(hipStreamSynchronize(m->ctx->stream));
        acc[q].resize(n);
        for (int x = 0; x < g.NX; ++x)
            for (int y = 0; y < g.NY; ++y)
                for (int z = 0; z < g.NZ; ++z)
                    acc[q][((size_t)x * g.NY + y) * g.NZ + z] = tmp[g.idx(z, y, x)];
    }
    return 0;
}
}  // namespace

int seis_grad_read(seis_model *m, float *grad) {
    g_err.clear();
    auto &G = m->grad;
    if (!G.ready) return fail("seis_grad_read: no gradient accumulated");
    std::vector<double> acc[3];
    const double *src[3] = {G.gM, G.gmu, G.grho};
    if (accum_read(m, src, 3, acc)) return 1;
    if (m->dft.nf) {
        // the DFT gradient is already in the parameterisation (no transf_grad)
        const size_t n = acc[0].size();
        for (int q = 0; q < 3; ++q)
            for (size_t i = 0; i < n; ++i) grad[q * n + i] = (float)acc[q][i];
        return 0;
    }
    const bool f0 = G.visco && (m->c.gradfix & FIX_VISCO_F0);
    if (G.grz) {
        // per-node terms through the staggered averages onto the cells
        std::vector<double> sh[3], rh[3];
        const double *ssrc[3] = {G.gsh[0], G.gsh[1], G.gsh[2]};
        const double *rsrc[3] = {G.grho, G.gry, G.grz};
        const int nq = m->c.ND == 3 ? 3 : 1;
        if (accum_read(m, ssrc, nq, sh)) return 1;
        if (m->c.ND == 3) {
            if (accum_read(m, rsrc, 3, rh)) return 1;
        } else {
            rh[0] = acc[2];
            if (accum_read(m, rsrc + 2, 1, rh + 2)) return 1;
        }
        chain_staggered(m->c, m->rp, f0 ? m->rp.muf0 : m->rp.mu, sh, rh, acc[1].data(),
                        acc[2].data());
    }
    transf_grad(m->c, m->rp, f0 ? m->rp.Mf0 : m->rp.M, f0 ? m->rp.muf0 : m->rp.mu,
                acc[0].data(), acc[1].data(), acc[2].data(), grad, m->c.gradfix & FIX_RHO_SCALE);
    return 0;
}

int seis_grad_read_device(seis_model *m, float *dgrad, float *dgtau) {
    g_err.clear();
    const size_t n = (size_t)m->g.NZ * m->g.NY * m->g.NX;
    hipStream_t st = m->ctx->stream;
    if (dgrad) {
        std::vector<float> h(3 * n);
        if (seis_grad_read(m, h.data())) return 1;
        HIP_TRY(hipMemcpyAsync(dgrad, h.data(), h.size() * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (dgtau) {
        std::vector<float> h(2 * n);
        if (seis_grad_read_tau(m, h.data())) return 1;
        HIP_TRY(hipMemcpyAsync(dgtau, h.data(), h.size() * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return 0;
}

int seis_grad_read_tau(seis_model *m, float *gtau) {
    g_err.clear();
    auto &G = m->grad;
    if (!G.ready || !G.gtp)
        return fail("seis_grad_read_tau: no tau gradient accumulated (viscoelastic runs, "
                    "param_type 0-2)");
    std::vector<double> acc[2];
    const double *src[2] = {G.gtp, G.gts};
    if (accum_read(m, src, 2, acc)) return 1;
    if (m->dft.nf) {
        const size_t n = acc[0].size();
        for (size_t i = 0; i < n; ++i) {
            gtau[i] = (float)acc[0][i];
            gtau[n + i] = (float)acc[1][i];
        }
        return 0;
    }
    // the accumulators hold dt^2/dh times dJ/dtau, as the moduli's hold
    // dt^2/dh times the derivative by the scaled moduli (transf_grad)
    const double k = (double)m->c.dh / ((double)m->c.dt * m->c.dt);
    const size_t n = acc[0].size();
    for (size_t i = 0; i < n; ++i) {
        gtau[i] = (float)(acc[0][i] * k);
        gtau[n + i] = (float)(acc[1][i] * k);
    }
    return 0;
}

int seis_hess_read(seis_model *m, float *hess) {
    g_err.clear();
    auto &G = m->grad;
    if (!G.ready || !G.hM) return fail("seis_hess_read: no Hessian accumulated (set Hout=1)");
    std::vector<double> acc[2];
    const double *src[2] = {G.hM, G.hmu};
    if (accum_read(m, src, 2, acc)) return 1;
    const bool f0 = G.visco && (m->c.gradfix & FIX_VISCO_F0);
    transf_hess(m->c, m->rp, f0 ? m->rp.Mf0 : m->rp.M, f0 ? m->rp.muf0 : m->rp.mu,
                acc[0].data(), acc[1].data(), hess);
    return 0;
}

int seis_gradient_shot(seis_model *m, int nsrc, const float *src_pos, const float *src, int nrec,
                       const float *rec_pos, const float *din, const float *res_in, float *out,
                       double *rms) {
    g_err.clear();
    if (!din && !res_in) return fail("seis_gradient_shot: give observed data or residuals");
    const int NT = m->c.NT, nout = (int)m->outn.size();
    std::vector<float> o((size_t)nout * nrec * NT), res(o.size(), 0.f);
    if (seis_forward_save(m, nsrc, src_pos, src, nrec, rec_pos, o.data())) return 1;
    double rr[2] = {0, 0};
    if (res_in) std::copy(res_in, res_in + res.size(), res.begin());
    else rr[0] = residuals(m->c, nout, nrec, o.data(), din, res.data(), &rr[1]);
    if (!res_in && m->dft.nf)
        rr[0] = dft_rms(m->c, nout, nrec, o.data(), din, m->dft.freqs, &rr[1]);
    res_scale(m->c, m->rp, m->outn, nrec, rec_pos, res.data(), m->c.gradfix & FIX_RES_SCALE);
    if (seis_adjoint(m, res.data())) return 1;
    if (out) std::copy(o.begin(), o.end(), out);
    if (rms) {
        rms[0] = rr[0];
        rms[1] = rr[1];
    }
    return 0;
}

// ---- batched serving -------------------------------------------------------

int seis_batch_create(seis_ctx *ctx, int max_slots, seis_batch **out) {
    g_err.clear();
    *out = nullptr;
    if (!ctx || max_slots <= 0 || max_slots > 4096) return fail("seis_batch_create: bad arguments");
    auto b = std::make_unique<seis_batch>();
    b->ctx = ctx;
    b->max_slots = max_slots;
    b->model.assign(max_slots, nullptr);
    b->host_slots.resize(max_slots);
    b->host_active.assign(max_slots, 0);
    b->t.assign(max_slots, 0);
    b->tmax.assign(max_slots, 0);
    void *p = nullptr;
    if (ctx->alloc(sizeof(BatchSlot) * max_slots, &p)) return 1;
    b->dslots = (BatchSlot *)p;
    if (ctx->alloc(sizeof(int) * max_slots * 2, &p)) return 1;
    b->dids = (int *)p;
    if (ctx->alloc(sizeof(unsigned) * (max_slots + 8) * 2, &p)) return 1;
    b->dtile0 = (unsigned *)p;
    for (int k = 0; k < 2; ++k) HIP_TRY(hipEventCreateWithFlags(&b->table_ev[k], hipEventDisableTiming));
    b->step_ev.resize(64);
    for (auto &e : b->step_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = b.release();
    return 0;
}

int seis_batch_destroy(seis_batch *b) {
    if (!b) return 0;
    (void)hipStreamSynchronize(b->ctx->stream);
    for (int i = 0; i < b->max_slots; ++i)
        if (b->model[i]) {
            b->model[i]->done_ev = nullptr;
            b->model[i]->in_batch = false;
        }
    for (int k = 0; k < 2; ++k) (void)hipEventDestroy(b->table_ev[k]);
    for (auto e : b->step_ev) (void)hipEventDestroy(e);
    b->ctx->release(b->dslots);
    b->ctx->release(b->dids);
    b->ctx->release(b->dtile0);
    delete b;
    return 0;
}

int seis_batch_add(seis_batch *b, seis_model *m, int *slot) {
    g_err.clear();
    if (!m->shot_bound) return fail("seis_batch_add: bind a shot first");
    if (m->ctx != b->ctx) return fail("seis_batch_add: the model lives in another context");
    const seis_csts &c = m->c;
    if (c.ND == 3 || c.L != 0 || c.freesurf || m->prec != P_FP32)
        return fail("seis_batch_add: batches take 2D FP32 jobs with L=0 and no free surface");
    int k = -1;
    for (int i = 0; i < b->max_slots; ++i)
        if (!b->model[i]) {
            k = i;
            break;
        }
    if (k < 0) return fail("seis_batch_add: the batch is full");
    BatchSlot sl;
    std::memset(&sl, 0, sizeof(sl));
    sl.a = m->a;
    sl.shot = m->shot;
    sl.active = 1;
    sl.nd_case = m->nd_case;
    sl.NT = c.NT;
    sl.tmax = tmax_of(m);
    sl.start = b->step;
    sl.nout = (int)m->out_codes.size();
    for (int q = 0; q < sl.nout && q < 10; ++q) sl.codes[q] = m->out_codes[q];
    // the job starts from zeroed wavefields and seismograms
    if (seis_launch_init(m)) return 1;
    b->host_slots[k] = sl;
    HIP_TRY(hipMemcpyAsync(b->dslots + k, &b->host_slots[k], sizeof(sl), hipMemcpyHostToDevice,
                           b->ctx->stream));
    b->model[k] = m;
    m->done_ev = nullptr;
    m->in_batch = true;
    b->t[k] = 0;
    b->tmax[k] = sl.tmax;
    b->dirty = true;
    *slot = k;
    return 0;
}

int seis_batch_remove(seis_batch *b, int slot) {
    g_err.clear();
    if (slot < 0 || slot >= b->max_slots || !b->model[slot]) return fail("seis_batch_remove: bad slot");
    b->host_active[slot] = 0;
    HIP_TRY(hipMemcpyAsync((char *)(b->dslots + slot) + offsetof(BatchSlot, active),
                           &b->host_active[slot], sizeof(int), hipMemcpyHostToDevice,
                           b->ctx->stream));
    b->model[slot]->done_ev = nullptr;
    b->model[slot]->in_batch = false;
    b->model[slot] = nullptr;
    b->dirty = true;
    return 0;
}

// slots still stepping, and the fewest steps any of them has left
int seis_batch_status(const seis_batch *b, int *nactive, int *min_left) {
    int n = 0, ml = 0;
    for (int i = 0; i < b->max_slots; ++i)
        if (b->model[i] && b->t[i] < b->tmax[i]) {
            const int left = b->tmax[i] - b->t[i];
            ml = n == 0 ? left : std::min(ml, left);
            ++n;
        }
    if (nactive) *nactive = n;
    if (min_left) *min_left = ml;
    return 0;
}

int seis_batch_slot_left(const seis_batch *b, int slot) {
    if (slot < 0 || slot >= b->max_slots || !b->model[slot]) return -1;
    return b->tmax[slot] - b->t[slot];
}

int seis_batch_step(seis_batch *b, int nsteps) {
    g_err.clear();
    hipStream_t st = b->ctx->stream;
    if (b->dirty) {
        // rebuild the class tables: slots grouped by (acoustic/elastic, FD order)
        std::map<std::pair<int, int>, std::vector<int>> groups;
        int maxrec = 0;
        for (int i = 0; i < b->max_slots; ++i) {
            seis_model *m = b->model[i];
            if (!m) continue;
            groups[{m->nd_case, m->g.H}].push_back(i);
            maxrec = std::max(maxrec, m->shot.nrec);
        }
        const int tb = b->tbuf ^ 1;
        HIP_TRY(hipEventSynchronize(b->table_ev[tb]));   // last launches reading this buffer
        b->tbuf = tb;
        int *dids = b->dids + (size_t)tb * b->max_slots;
        unsigned *dtile0 = b->dtile0 + (size_t)tb * (b->max_slots + 8);
        std::vector<int> &ids = b->host_ids[tb];
        std::vector<unsigned> &tile0 = b->host_tile0[tb];
        ids.clear();
        tile0.clear();
        unsigned acc = 0;
        // tiles per block of the batch passes (SEIS_BATCH_TILES, 1..16)
        static const int bt = [] {
            const char *v = getenv("SEIS_BATCH_TILES");
            const int n = v && *v ? atoi(v) : kBatchTiles;
            return n < 1 ? 1 : n > 16 ? 16 : n;
        }();
        b->table.tiles = bt;
        for (auto &kv : groups)
            for (int i : kv.second) {
                const Layout &g = b->model[i]->g;
                ids.push_back(i);
                tile0.push_back(acc);
                const unsigned tiles = (unsigned)((g.NZ + 63) / 64) * (unsigned)((g.NX + 3) / 4);
                acc += (tiles + bt - 1) / bt;
            }
        tile0.push_back(acc);
        b->table.n = (int)ids.size();
        b->table.nblocks = acc;
        b->table.ids = dids;
        b->table.block0 = dtile0;
        if (!ids.empty())
            HIP_TRY(hipMemcpyAsync(dids, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dtile0, tile0.data(), tile0.size() * 4, hipMemcpyHostToDevice, st));
        b->maxrec = maxrec;
        b->dirty = false;
    }
    int hi = 0;
    for (int i = 0; i < b->max_slots; ++i)
        if (b->model[i]) hi = i + 1;
    launch_batch_steps(b->dslots, hi, b->table, b->maxrec, b->step, nsteps, st);
    b->step += nsteps;
    HIP_TRY(hipEventRecord(b->table_ev[b->tbuf], st));
    hipEvent_t ev = b->step_ev[b->step_ev_next++ % b->step_ev.size()];
    HIP_TRY(hipEventRecord(ev, st));
    for (int i = 0; i < b->max_slots; ++i)
        if (b->model[i] && b->t[i] < b->tmax[i]) {
            b->t[i] = std::min(b->tmax[i], b->t[i] + nsteps);
            if (b->t[i] == b->tmax[i]) b->model[i]->done_ev = ev;
        }
    HIP_TRY(hipGetLastError());
    return 0;
}

int seis_last_timing(const seis_model *m, double *ms_v, double *ms_s, double *ms_loop) {
    if (ms_v) *ms_v = m->ms_v;
    if (ms_s) *ms_s = m->ms_s;
    if (ms_loop) *ms_loop = m->ms_loop;
    return 0;
}

}  // extern "C"
