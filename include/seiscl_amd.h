/*
 * seiscl_amd.h -- C-ABI of the MI355X-native SeisCL time-domain FD engine.
 *
 * This library replaces the device runtime of the reference (Init_OpenCL /
 * clprogram / clbuf, src/F.h:124-189, 547, 566) and its per-shot time loop
 * (time_stepping, src/time_stepping.c:692-1012) with hand-written gfx950 HIP
 * kernels.  Everything crossing this boundary is plain C: opaque handles,
 * pointers, sizes and int status codes (0 = success).  No call writes to
 * stderr; the text of the last error on the calling thread is returned by
 * seis_last_error().
 *
 * Array layouts at this boundary are the reference's host layouts:
 *   - model parameters: N[0]=NZ (fastest), then NY, then NX, unpadded
 *     (what SeisCL.py writes after np.transpose, SeisCL/SeisCL.py:747-763);
 *   - src_pos [nsrc][5] = x, y, z, srcid, type   (src/read_hdf5.c:307-459);
 *   - src     [nsrc][NT];
 *   - rec_pos [nrec][8] = x, y, z, srcid, recid, type, -, -;
 *   - seismograms [nout][nrec][NT] (src/time_stepping.c:25-90), outputs in
 *     the order of seis_out_name();
 *   - gradients: unpadded like the parameters.
 */
#ifndef SEISCL_AMD_H
#define SEISCL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEIS_MAXL 8

/* Modeling constants.  Field names, meaning and defaults follow the HDF5
 * constants file read by read_csts (src/read_hdf5.c:461-552) and the csts
 * list of SeisCL.py (SeisCL/SeisCL.py:23-30).  Fill with seis_csts_default()
 * then override. */
typedef struct seis_csts {
    int ND;             /* 2: 2D P-SV, 3: 3D, 22: 2D acoustic                   */
    int N[3];           /* [NZ, NX] (2D) or [NZ, NY, NX] (3D)                    */
    int NT;             /* time samples                                           */
    float dt, dh;
    int FDORDER;        /* 2, 4, 6, 8, 10, 12                                     */
    int MAXRELERROR;    /* 0 Taylor, 1..4 Holberg                                 */
    int freesurf;       /* 1: free surface on top (z=0)                           */
    int abs_type;       /* 1: CPML, 2: Cerjan taper, 0: none                     */
    int nab;            /* absorbing width in cells                               */
    float abpc, VPPML, NPOWER, FPML, K_MAX_CPML;
    int L;              /* Maxwell bodies (0 = elastic)                           */
    float f0;           /* reference frequency of the phase velocity              */
    float FL[SEIS_MAXL];/* relaxation frequencies                                 */
    int seisout;        /* 1 velocities, 2 pressure, 3 stresses, 4 all           */
    int param_type;     /* 0 vp,vs,rho  1 M,mu,rho  2 Ip,Is,rho  3 complex v      */
    int tmax;           /* forward steps, 0 -> NT   (src/read_hdf5.c:523-526)    */
    int tmin;           /* first adjoint step bound (src/time_stepping.c:869)     */
    int gradout;        /* 1: keep what the gradient needs                        */
    int FP16;           /* 0: FP32; 1: FP32 in the scaled units of 2; 2: FP16    */
                        /* storage, FP32 compute; 64: FP64                        */
    int gradfix;        /* gradient corrections of reference defects (bit set):   */
                        /*  1 res_scale reads the buoyancy at the receiver's own  */
                        /*    cell (src/residuals.c:487-491 uses x three times);  */
                        /*  2 the adjoint run has its own zeroed CPML memory (the */
                        /*    reference's adjoint kernels bind the forward one);  */
                        /*  4 the direct density term is scaled by dh^2/dt^2      */
                        /*    (src/calc_grad.c:1005-1013 omits it);               */
                        /*  8 the 3D adjoint free surface reads syy (the          */
                        /*    reference reads sxx twice, src/surface3D_adj.cl:69);*/
                        /* 16 viscoelastic gradients use the moduli at f0 for the */
                        /*    stress-to-strain conversion and the chain rule (the */
                        /*    reference uses the relaxed ones, src/update_adjs2D. */
                        /*    cl:441-451, src/calc_grad.c:1020-1056);             */
                        /* 32 vs and rho gradients through the staggered averages */
                        /*    of the transform (the reference gives each node's   */
                        /*    term to the cell of the same index); not in 223     */
                        /* 64 normalised-misfit adjoint sources (scalermsnorm)    */
                        /*    scale the trace by o.d, the derivative of           */
                        /*    1/2 |o/|o| - d/|d||^2 (src/residuals.c:331-337 use  */
                        /*    1 - o.d);                                           */
                        /* 128 the mu Hessian subtracts the Hessian's own M term  */
                        /*    (the 2D reference subtracts the gradient's,         */
                        /*    src/update_adjs2D.cl:456).                          */
                        /* seis_csts_default sets 223; 0 reproduces the reference.*/
    /* misfit options of var_res_raw (src/residuals.c:58-450), read by       */
    /* seis_residuals and seis_gradient_shot:                                */
    int restype;        /* 0 L2 residuals; 1 the data themselves (rtm_res,      */
                        /*   src/residuals.c:24-55)                              */
    int scalerms;       /* traces scaled by their rms, misfit normalised        */
    int scaleshot;      /* traces scaled by the rms of the whole shot           */
    int scalermsnorm;   /* normalised misfit without per-trace scaling          */
    float fmin, fmax;   /* Butterworth band-pass (order 6, zero phase) of data  */
                        /*   and seismograms before the misfit, 0: off          */
                        /*   (src/butterworth.c:24-86)                          */
    int Hout;           /* 1: accumulate the approximate Hessian of M and mu    */
                        /*   with the gradient (src/update_adjs2D.cl:453-456)   */
} seis_csts;

typedef struct seis_ctx seis_ctx;       /* one GPU: device, streams, allocator   */
typedef struct seis_model seis_model;   /* one modeling case resident on a GPU   */
typedef struct seis_batch seis_batch;   /* modeling jobs stepped together        */

/* ---- errors ------------------------------------------------------------- */
const char *seis_last_error(void);
void seis_csts_default(seis_csts *c);    /* SeisCL.py:36-83 defaults             */
size_t seis_csts_size(void);             /* sizeof(seis_csts), for FFI checks    */

/* ---- device context: replaces Init_CUDA device discovery and the two      */
/* queues per device (src/Init_OpenCL.c:24-323, 274-287)                      */
int seis_ctx_create(int device_ordinal, seis_ctx **ctx);
int seis_ctx_destroy(seis_ctx *ctx);
int seis_ctx_sync(seis_ctx *ctx);                   /* WAITQUEUE (src/F.h:634)  */
size_t seis_ctx_footprint(const seis_ctx *ctx);     /* device bytes allocated   */

/* ---- buffers: replace clbuf_create / clbuf_send / clbuf_read              */
/* (src/clbuf.c:23-281).  Transfers are ordered on the compute stream.        */
/* Freed device memory stays with the context and is reused by later         */
/* allocations (no device-wide synchronisation while jobs come and go); it   */
/* returns to the system at seis_ctx_destroy.  seis_ctx_footprint counts the */
/* bytes in use.                                                              */
int seis_buf_create(seis_ctx *ctx, size_t bytes, void **dptr);
int seis_buf_free(seis_ctx *ctx, void *dptr);
int seis_buf_send(seis_ctx *ctx, void *dptr, const void *host, size_t bytes);
int seis_buf_read(seis_ctx *ctx, void *host, const void *dptr, size_t bytes);

/* ---- model: replaces assign_modeling_case + Init_cst/Init_model +          */
/* the buffer/program set-up of Init_CUDA (src/assign_modeling_case.c:871-   */
/* 1326, src/Init_model.c:40-97, src/Init_OpenCL.c:326-1217).  Parameters are */
/* those named by param_type (vp|M|Ip, vs|mu|Is, rho, taup, taus); pointers  */
/* the case does not use may be NULL.  Fails on an unstable dt                */
/* (check_stability, src/assign_modeling_case.c:730-827).                     */
int seis_model_create(seis_ctx *ctx, const seis_csts *c, const float *p0,
                      const float *p1, const float *rho, const float *taup,
                      const float *taus, seis_model **model);
int seis_model_destroy(seis_model *m);
/* Replace the material parameters of an existing model (same grid).          */
int seis_model_set_params(seis_model *m, const float *p0, const float *p1,
                          const float *rho, const float *taup, const float *taus);
/* The same with the parameters in device memory of the model's GPU (e.g. */
/* torch tensors, unpadded in the layout above), ready when called.  The    */
/* transforms run on the host as for seis_model_set_params (Init_model's    */
/* host side, src/Init_model.c:40-97), after one device-to-host copy.         */
int seis_model_set_params_device(seis_model *m, const float *dp0, const float *dp1,
                                 const float *drho, const float *dtaup,
                                 const float *dtaus);
size_t seis_model_footprint(const seis_model *m);

/* number / names of seismogram outputs for the model's seisout              */
/* (src/assign_modeling_case.c:1244-1326; SeisCL.py to_load_names :376-408)  */
int seis_out_count(const seis_model *m);
const char *seis_out_name(const seis_model *m, int i);

/* ---- per-shot runtime: the launch sequence of time_stepping.               */
/* seis_shot_bind = initialize_forward H2D of src/src_pos/rec_pos            */
/* (src/time_stepping.c:553-621); the step launches replace prog_launch of   */
/* sources / update_v / update_s / freesurface / varsout                     */
/* (src/time_stepping.c:785-816, src/clprogram.c:1004-1071).                  */
int seis_shot_bind(seis_model *m, int nsrc, const float *src_pos,
                   const float *src, int nrec, const float *rec_pos);
int seis_launch_init(seis_model *m);                 /* vars_init + varsoutinit */
int seis_launch_sources(seis_model *m, int nt, int pdir);
int seis_launch_update_v(seis_model *m);
int seis_launch_update_s(seis_model *m);
int seis_launch_surface(seis_model *m);
int seis_launch_varsout(seis_model *m, int nt);
/* D2H of the bound shot's seismograms, [nout][nrec][NT] (reduce_seis,       */
/* src/time_stepping.c:25-90).                                                */
int seis_shot_read(seis_model *m, float *out);
/* The same into device memory of the model's GPU, [nout][nrec][NT] floats  */
/* (reduce_seis, src/time_stepping.c:25-90, without leaving the GPU).         */
int seis_shot_read_device(seis_model *m, float *dout);
/* The same in double, for models running in FP64 (FP16=64), whose receivers */
/* record in double.  No counterpart in the reference, which has no FP64     */
/* mode; finite-difference checks of gradients need the digits.             */
int seis_shot_read_f64(seis_model *m, double *out);

/* Run forward steps t0 .. t0+nsteps-1 of the bound shot (sources, update_v, */
/* update_s, free surface, varsout), waiting for completion; the device time */
/* of the run is returned by seis_last_timing.                                */
int seis_run_steps(seis_model *m, int t0, int nsteps);

/* Whole forward shot: bind + init + NT steps + read (the forward half of    */
/* time_stepping, src/time_stepping.c:747-848).                              */
int seis_forward_shot(seis_model *m, int nsrc, const float *src_pos,
                      const float *src, int nrec, const float *rec_pos,
                      float *out);

/* Size of the fixed device area seis_forward_shot_static works in (one per */
/* GPU, held until the process ends).                                         */
#define SEIS_STATIC_SHOT_BYTES ((size_t)64 << 20)

/* seis_forward_shot without allocating device memory.  No counterpart in   */
/* the reference: for callers that cannot call hipMalloc (code under HIP    */
/* graph capture, serving loops, harnesses that forbid allocation).  The     */
/* shot's source and receiver tables, its source samples, the seismogram    */
/* recording and its transpose live in a static area of                      */
/* SEIS_STATIC_SHOT_BYTES per GPU; a shot needing more (about                 */
/* 8 x NT x nrec x outputs bytes, twice that in FP64 runs, plus 4 x NT x nsrc) */
/* is refused with an error.  Inputs, outputs and results are those of      */
/* seis_forward_shot; the model's wavefields must already exist              */
/* (seis_model_create).  Afterwards no shot is bound to the model.  Every    */
/* call on one GPU shares that GPU's area: calls that overlap on one GPU     */
/* race.  Models in a serving batch are refused.                             */
int seis_forward_shot_static(seis_model *m, int nsrc, const float *src_pos,
                             const float *src, int nrec, const float *rec_pos,
                             float *out);

/* Copy one wavefield ("vx", "sxx", ...) to the host in the reference's     */
/* padded layout, Pi(N+FDORDER) floats (used by movout, src/time_stepping.c: */
/* 272-350, and by parity tests).                                              */
int seis_read_field(seis_model *m, const char *name, float *host);

/* ---- gradient by backpropagation (BACK_PROP_TYPE=1), ND 2/3, with or    */
/* without a free surface (adjoint surface: src/surface2D_adj.cl,            */
/* src/surface3D_adj.cl).  Elastic (L=0, FP32): the forward run saves an     */
/* FDOH-thick boundary shell of every wavefield at every step in device      */
/* memory (save_bnd, src/time_stepping.c:352-395, kept resident instead of   */
/* the reference's pinned host copy); the adjoint run re-injects it while    */
/* running the forward wavefield backwards (inject_bnd, :397-424).           */
/* Viscoelastic (L=1..4; FP32, FP64, FP16 storage or scaled units):          */
/* running the memory variables backwards is unstable, so the forward run    */
/* keeps full-state checkpoints in device memory and the adjoint recomputes  */
/* the forward run segment by segment (one or two checkpoint levels, chosen  */
/* to fit a memory budget), keeping the post-source states it reads.  Both   */
/* accumulate the gradient in double (update_adjs / update_adjv,             */
/* src/update_adj*.cl).                                                      */
/* Zero the gradient accumulators (kernel_gradinit, src/automatic_kernels.c: */
/* 840-930; src/time_stepping.c:725-729).                                     */
int seis_grad_zero(seis_model *m);
/* Device bytes the viscoelastic gradient may hold in checkpoints and the   */
/* recomputed history (0, the default: 90% of the free device memory at    */
/* seis_forward_save).  No counterpart in the reference, whose viscoelastic */
/* adjoint backpropagates.  seis_grad_plan reports the plan chosen: levels  */
/* (1..4; each level costs one recomputation of the forward run), K[4] the  */
/* spacing of each level's checkpoints (K[levels-1]: steps of recomputed    */
/* history), n the ratio of consecutive spacings, n0 the checkpoints of the */
/* forward run, and the bytes the plan holds.                               */
int seis_grad_set_memory(seis_model *m, size_t bytes);
int seis_grad_plan(const seis_model *m, int *levels, int *K, int *n, int *n0,
                   size_t *bytes);
/* Bind a shot and run it forward with boundary saving; out as             */
/* seis_forward_shot (may be NULL).                                          */
int seis_forward_save(seis_model *m, int nsrc, const float *src_pos,
                      const float *src, int nrec, const float *rec_pos,
                      float *out);
/* The saved forward run of the bound shot as one host buffer, and back:    */
/* the device side of checkpoint_d2h / checkpoint_h2d (src/time_stepping.c: */
/* 92-270) for the two-pass inputres flow (:731-745, 851-859).  Elastic     */
/* runs save the final wavefields and CPML memory with the boundary shell   */
/* of every step; viscoelastic runs save the forward run's full-state       */
/* checkpoints.  The buffer starts with a header of this library's own      */
/* (magic "SEISCKP1", kind, steps, sizes); seis_grad_saved_write checks it  */
/* against the model and its checkpoint plan, needs the same shot bound     */
/* (seis_shot_bind) and makes seis_adjoint run without a forward run.       */
int seis_grad_saved_size(seis_model *m, size_t *bytes);
int seis_grad_saved_read(seis_model *m, float *host);
int seis_grad_saved_write(seis_model *m, const float *host);
/* Residuals (adjoint sources before res_scale) of a shot, [nout][nrec][NT]: */
/* var_res_raw (src/residuals.c:58-450) with the model's misfit options --    */
/* Butterworth band-pass (fmin, fmax), rms scalings (scalerms, scaleshot,     */
/* scalermsnorm) -- or the data themselves with restype 1 (rtm_res,           */
/* src/residuals.c:24-55); rms = sum (din-out)^2, rmsnorm = sum din^2 of the  */
/* filtered and scaled traces.  Host only.                                    */
int seis_residuals(const seis_model *m, int nrec, const float *out,
                   const float *din, float *res, double *rms, double *rmsnorm);
/* The same for constants c and nout outputs, without a model (host only, no */
/* device needed).                                                            */
int seis_misfit(const seis_csts *c, int nout, int nrec, const float *out,
                const float *din, float *res, double *rms, double *rmsnorm);
/* Scale adjoint sources in place (res_scale, src/residuals.c:452-570).      */
int seis_res_scale(const seis_model *m, int nrec, const float *rec_pos,
                   float *res);
/* Adjoint run of the shot bound by seis_forward_save with scaled adjoint   */
/* sources res [nout][nrec][NT]: steps tmax-1 .. tmin+1 (src/time_stepping.c:*/
/* 869-940); adds this shot's gradient to the accumulators.                 */
int seis_adjoint(seis_model *m, const float *res);
/* seis_adjoint with the adjoint sources in device memory ([nout][nrec][NT],*/
/* ready when called); res_scale != 0 applies res_scale (as seis_res_scale  */
/* with the bound shot's receivers) on the device first.                     */
int seis_adjoint_device(seis_model *m, const float *dres, int res_scale);
/* seis_adjoint recording the movie of the gradient run (movout with gradout,*/
/* src/time_stepping.c:926-929, frame index :317-320): after the adjoint    */
/* step t with t % movout == 0, frame (t+1)/movout - 1 of each wavefield    */
/* vars[v] ("vx", "sxx", ...) holds the backpropagated forward wavefield.  */
/* Viscoelastic runs read the forward state from a recomputed history and  */
/* record the forward wavefield at the start of step t instead, the state  */
/* backpropagation reconstructs there.  frames: [nvar][NT/movout][Pi N],    */
/* interior cells in the parameter layout, physical units; frames no step  */
/* writes are zero.                                                         */
int seis_adjoint_movie(seis_model *m, const float *res, int movout, int nvar,
                       const char *const *vars, float *frames);
/* Gradients [3][Pi N] in param_type order (vp|M|Ip, vs|mu|Is, rho), in the */
/* parameter layout (transf_grad, src/calc_grad.c:975-1110).                 */
int seis_grad_read(seis_model *m, float *grad);
/* Gradients of tau_p and tau_s, [2][Pi N] in the parameter layout, of a    */
/* viscoelastic model (L > 0, param_type 0-2), at fixed wave speeds (the    */
/* other parameters of param_type held).  In the reference they exist only */
/* in the DFT path (grad_coefvisc_0, src/calc_grad.c:111-152, written as    */
/* gradtaup/gradtaus by src/writehdf5.c:261-283); its time-domain kernels   */
/* declare and never write them (src/update_adjs3D.cl:66).  Here they are  */
/* accumulated in double by the time-domain adjoint kernels from the whole */
/* memory term of each step.                                                */
int seis_grad_read_tau(seis_model *m, float *gtau);
/* seis_grad_read and (dgtau not NULL) seis_grad_read_tau into device       */
/* memory of the model's GPU (gradout of src/time_stepping.c:948-1003 for a  */
/* caller whose tensors live there); the transform runs on the host.         */
int seis_grad_read_device(seis_model *m, float *dgrad, float *dgtau);
/* The approximate Hessian accumulated with the gradient when Hout=1,       */
/* [3][Pi N] in the same order and transform (H<name> of the gradient file,  */
/* src/writehdf5.c:275-283, src/calc_grad.c:1028-1094).                      */
int seis_hess_read(seis_model *m, float *hess);
/* DFT-domain gradient (back_prop_type 2; src/calc_grad.c:377-968, the     */
/* savefreqs kernel src/automatic_kernels.c:1030-1128, sampling             */
/* src/assign_modeling_case.c:549-570, src/time_stepping.c:658-672,767-780, */
/* 869-965): the model's next gradient runs keep the DFT of every wavefield */
/* and memory variable at the nfreqs frequencies freqs (Hz, gradfreqs)     */
/* every DTNYQ = ceil(0.0156/(fmax dt)) steps of the forward run, propagate */
/* the adjoint wavefield alone and keep its DFT, and form the gradient of  */
/* the parameterisation (vp|M|Ip, vs|mu|Is, rho, and taup, taus for L>0)    */
/* from the two sets of spectra; seis_grad_read / seis_grad_read_tau then  */
/* return it as accumulated (no transf_grad), and seis_residuals /         */
/* seis_gradient_shot report the rms at these frequencies                 */
/* (src/residuals.c:364-403).  nfreqs = 0 returns to backpropagation; zero */
/* the accumulators (seis_grad_zero) after switching.  FP32 models, ND 2   */
/* or 3, L2 misfit without trace scaling, no Hessian, no RTM.              */
int seis_grad_dft(seis_model *m, int nfreqs, const float *freqs);
/* One shot of gradient: forward with boundary saving, residuals against    */
/* din (or the given adjoint sources res_in, the inputres mode), res_scale, */
/* adjoint.  out: forward seismograms (may be NULL); rms: [rms, rmsnorm].   */
int seis_gradient_shot(seis_model *m, int nsrc, const float *src_pos,
                       const float *src, int nrec, const float *rec_pos,
                       const float *din, const float *res_in, float *out,
                       double *rms);

/* ---- batched serving (no counterpart in the reference, which runs one    */
/* shot at a time, src/time_stepping.c:747): up to max_slots modeling jobs, */
/* each a model with a bound shot, advance together, one launch per kernel  */
/* class and time step for all of them; a job's results are identical      */
/* whether it runs alone or in a batch.  2D, FP32, L=0, no free surface.     */
int seis_batch_create(seis_ctx *ctx, int max_slots, seis_batch **batch);
int seis_batch_destroy(seis_batch *b);
/* Add a job (zeroes its wavefields; it starts at step 0); *slot receives   */
/* its slot.  Fails when the batch is full.                                   */
int seis_batch_add(seis_batch *b, seis_model *m, int *slot);
/* Free a slot (read its seismograms first with seis_shot_read).             */
int seis_batch_remove(seis_batch *b, int slot);
/* Advance every job by nsteps steps (jobs at their last step stay idle).    */
/* Asynchronous: seis_shot_read or seis_ctx_sync wait for it.                */
int seis_batch_step(seis_batch *b, int nsteps);
/* Jobs still stepping and the fewest steps any of them has left.           */
int seis_batch_status(const seis_batch *b, int *nactive, int *min_left);
/* Steps a slot still has to run (0: done), -1 for a free slot.             */
int seis_batch_slot_left(const seis_batch *b, int slot);

/* Record HIP events around every update_v / update_s launch of the next    */
/* forward runs (costs a few microseconds per step).                          */
int seis_model_set_profiling(seis_model *m, int on);

/* Last timed loop: device milliseconds spent in the update kernels and in  */
/* the whole step loop (HIP events on the compute stream).  After            */
/* seis_run_steps: the forward update_v / update_s; after seis_adjoint: its  */
/* adjoint velocity / stress kernels (the whole adjoint loop includes the    */
/* recomputations of a viscoelastic model).  The kernel                      */
/* times need seis_model_set_profiling(m, 1) before the loop.                */
int seis_last_timing(const seis_model *m, double *ms_update_v,
                     double *ms_update_s, double *ms_loop);

#ifdef __cplusplus
}
#endif
#endif
