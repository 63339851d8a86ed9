/*
 * gft.h -- C ABI of libgft.so, the MI355X (gfx950) evolver for real-time
 * lattice SU(2)xU(1)-Higgs dynamics in temporal gauge.
 *
 * The reference evolves one lattice through per-site virtual C++ calls
 * (Potential / Gradient / WilsonLoop / Model, driven by Lattice::evolve(),
 * Core/Source/Lattice.cpp:755-876).  Per-site calls cannot cross to a GPU,
 * so this ABI cuts the path at whole-step granularity: one call advances the
 * whole lattice N leapfrog steps, with the reference's Energy and
 * GaugeCondition analysers folded into the same call.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative GFT_E* code on
 *     failure; gft_last_error() gives the message.  No C++ exception crosses
 *     this boundary (the reference throws std::runtime_error from its
 *     constructors, Core/Source/Model.cpp:38,85,114; the C++ host wrappers in
 *     gftevolver_amd/host rethrow with the same text).
 *   - The caller owns every host array; the library copies in and out.
 *   - A context owns its device buffers and one HIP stream.  A context is not
 *     thread-safe (neither is the reference: Model keeps per-site scratch,
 *     Core/Header/Model.hpp:255-262).
 *   - Two-slot arrays use the reference layout
 *       scalar[2][nx][ny][nz][Ns], vector[2][nx][ny][nz][Nv]
 *     (Core/Source/Lattice.cpp:358-367, 658-671) with Nv = 15 (three
 *     directions of [y, c0, c1, c2, c3]: U(1) phase then SU(2) quaternion),
 *     or Nv = 12 ([y, w1, w2, w3], generator form) when gft_desc.gen_rep is set.
 *     At step index k the current slot is (k+1)%2 and new values go to slot
 *     k%2 (Lattice.cpp:773-774).
 */
#ifndef GFT_H
#define GFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- models (Core/Config/Model.cfg potential/gradient/wilson-loop ids) ---- */
#define GFT_MODEL_SM    0  /* SO_N potential, Gradients::StandardModel, WilsonLoops::StandardModel; Ns=4 */
#define GFT_MODEL_2HDM  1  /* Double_SO_N, Gradients::TwoHDM, WilsonLoops::TwoHDM with gaugeNum 4; Ns=8 */
#define GFT_MODEL_GLOBAL 2 /* global O(N) model: SO_N (or Double_SO_N) potential, Gradients::Global,
                              null Wilson loop; Ns = num_scalar, no links, no Gauss law; two-slot forms */

/* ---- state forms ---- */
#define GFT_FORM_TWO_SLOT_F64 0  /* reference two-slot layout, FP64 storage (FP64 GPU oracle)            */
#define GFT_FORM_TWO_SLOT_F32 1  /* reference two-slot layout, FP32 storage (the reference's own storage) */
#define GFT_FORM_MOMENTUM_F32 2  /* structure-of-arrays kick/drift form, FP32 bulk (the fast path): FP32   */
                                 /* arithmetic and neighbour traffic; phi stored in FP32 as its          */
                                 /* difference from a constant reference doublet (the mean at import),   */
                                 /* P and E as FP32 + 4-bit residual in units of the FP32 value's ulp    */
                                 /* (about 28 bits)                                                      */
#define GFT_FORM_MOMENTUM_F64 3  /* structure-of-arrays kick/drift form, FP64 (plain FP64 GPU evolver)    */
#define GFT_FORM_MOMENTUM_F32_PLAIN 4 /* as MOMENTUM_F32 with plain FP32 P and E (no residuals): 5% fewer  */
                                 /* bytes, Gauss-law drift of FP32 rounding (~1e-6 after 1e4 steps)       */

/* ---- error codes ---- */
#define GFT_OK          0
#define GFT_EINVAL     -1  /* invalid argument or configuration                */
#define GFT_EHIP       -2  /* HIP runtime error                                */
#define GFT_ENOMEM     -3  /* device allocation failed                         */
#define GFT_EDOMAIN    -4  /* |R| >= 1 in a link update (asin/sqrt domain)      */
#define GFT_ENODEV     -5  /* no usable gfx950 device                          */

typedef struct gft_desc {
    int      model;            /* GFT_MODEL_*                                                  */
    int      form;             /* GFT_FORM_*                                                   */
    int      nx, ny, nz;       /* lattice size (Lattice.cfg line 11)                           */
    double   dx, dy, dz, dt;   /* spacings and time step (Lattice.cfg lines 12, 14)            */
    double   g, gp;            /* isospin / hypercharge couplings (WilsonLoops StandardModel.cfg) */
    double   div_tol;          /* division-by-zero tolerance for 1/g^2 (same cfg), 1e-9        */
    /* SO_N (Potentials/Source/SO_N.cpp:57-87): V = lambda/4 (|f|^2 - eta_sqr)^2              */
    double   lambda, eta_sqr;
    /* Double_SO_N (Potentials/Source/Double_SO_N.cpp:157-368)                                 */
    double   m1, m2, l1, l2, l3;
    int      break_to_son, break_to_uhalfn;
    double   m12r, l4p5, l6r, l7r;
    double   m12i, l4m5, l5i, l6i, l7i;
    int      pot_norm;         /* add pot_offset to the potential-energy density (energy only)  */
    double   pot_offset;
    /* damping (Core/Source/Model.cpp:241-254, 345): gamma for time_iter < nt_damped          */
    float    damping;
    unsigned nt_damped;
    /* Fixed faces (Lattice.cfg face types, Lattice.cpp:217-356; Fixed.cpp:159-251): bit 2a set
     * fixes the low face of axis a (x, y, z), bit 2a+1 the high face; a face's pair must match
     * (the reference requires periodic faces in opposite pairs).  A site in the one-site shell
     * of a fixed face is frozen: never updated (both time slots keep its initial values, which
     * must be equal there), still read by its neighbours, and left out of the energy and Gauss
     * sums.  0 = fully periodic.                                                              */
    int      fixed_faces;
    /* SU(2) link representation of the two-slot arrays (WilsonLoops/Source/StandardModel.cpp:
     * 78-157, 195-210; Lattice.cfg vector components 5 or 4; TwoHDM.cfg generator flag):
     * 0 = quaternion, vector[2][N][15] = 3 x [y, c0, c1, c2, c3];
     * 1 = generator,  vector[2][N][12] = 3 x [y, w1, w2, w3], U = exp(i w.sigma) =
     *     cos|w| + i sin|w| (w/|w|).sigma (w = 0 is the identity).
     * The two-slot forms step the stored form as the reference does (convert to a
     * quaternion, multiply, convert back); the momentum forms convert at import/export. */
    int      gen_rep;
    /* Gauge group of the 2HDM (WilsonLoops/Source/TwoHDM.cpp:215-242, gaugeNum): 0 or 4 =
     * SU(2) x U(1); 5 adds a Higgs-family U(1) (one more phase y' per link, opposite charge
     * on the two doublets); 7 adds a Higgs-family SU(2) acting on the doublet index.  Link
     * reals per direction: 4 -> 5 | 4 (quaternion | generator form), 5 -> 6 | 5
     * ([.., y']), 7 -> 9 | 7 ([.., c0'..c3'] | [.., w1'..w3']).  5 and 7 need GFT_MODEL_2HDM;
     * every form carries them (the momentum forms store the family phases or quaternions
     * and their electric fields beside the SU(2) x U(1) ones).  gpp is the Higgs-family
     * coupling g'' (TwoHDM.cfg).                                                          */
    int      gauge_num;
    double   gpp;
    /* GFT_MODEL_GLOBAL only (Gradients/Source/Global.cpp; Potentials/Source/SO_N.cpp takes any
     * component count): num_scalar = N (1..8); potential 0 or 1 = SO_N, 2 = Double_SO_N
     * (needs N = 8).  The vector arrays of the two-slot calls are ignored (may be NULL).  */
    int      num_scalar;
    int      potential;
} gft_desc;

/* Most Gauss-law generators of any gauge group (gaugeNum 7). */
#define GFT_MAX_EQ 7

/* Observables of one measured step.  energy[] columns follow the reference's
 * global_quantities.dat order (Analysers/Source/Energy.cpp:169-189) and are
 * already multiplied by dx*dy*dz; gauss_* follow global_quantities_gauge.dat
 * (Analysers/Source/GaugeCondition.cpp:167-194).  All sums are FP64
 * compensated (Neumaier) over FP64 per-site values. */
typedef struct gft_obs {
    unsigned time_iter;      /* step index the row belongs to                                  */
    unsigned has_energy;     /* energy[] valid (measured before the update of this step)       */
    unsigned has_gauss;      /* gauss_*[] valid (measured after the update of this step)       */
    unsigned num_eq;         /* number of Gauss generators: gaugeNum (4, 5 or 7)               */
    double   energy[6];      /* total, potential, gradient, kinetic, magnetic, electric        */
    double   gauss_int[GFT_MAX_EQ];   /* sum_x |G^a(x)| dx dy dz, a = Y, 1, 2, 3, family..     */
    double   gauss_max[GFT_MAX_EQ];   /* max_x |G^a(x)|                                        */
    double   gauss_scale[GFT_MAX_EQ]; /* scale of the relative violation: sum_x of the magnitudes of the
                                terms that cancel in G^a, sum_i (|E term at x| + |E term at x-i|)
                                / dx_i^2 + g_a^2 |rho^a|, times dx dy dz                      */
} gft_obs;

typedef struct gft_ctx gft_ctx;

/* Lifetime --------------------------------------------------------------- */
int         gft_create(const gft_desc* desc, gft_ctx** out);
void        gft_destroy(gft_ctx* ctx);
const char* gft_last_error(const gft_ctx* ctx);   /* ctx may be NULL: last create error */
int         gft_device_info(char* name, size_t name_len, int* cu_count);

/* State transfer (reference two-slot layout, see above).  time_iter is the
 * reference step counter of the next step to run; it fixes which slot is
 * current. */
int gft_import_two_slot_f32(gft_ctx* ctx, const float* scalar, const float* vec, unsigned time_iter);
int gft_import_two_slot_f64(gft_ctx* ctx, const double* scalar, const double* vec, unsigned time_iter);
int gft_export_two_slot_f32(gft_ctx* ctx, float* scalar, float* vec, unsigned* time_iter);
int gft_export_two_slot_f64(gft_ctx* ctx, double* scalar, double* vec, unsigned* time_iter);

/* Checkpoints (the reference dumps fields as 6-digit text with no reader,
 * Analysers/Source/OutputFields.cpp:63-133; binary resume is this library's
 * own).  A checkpoint holds every device array of the state -- for the momentum
 * forms both ping-pong buffers, P, E and the compensated storage's residual
 * words -- with the step counter, the current buffer and the reference
 * doublets, in a header-tagged host buffer of gft_checkpoint_size bytes.
 * gft_restore accepts only a checkpoint of a context with the same form, model,
 * lattice, gauge group, link representation and fixed faces, and a restored
 * context evolves on bitwise as the checkpointed one would have. */
int gft_checkpoint_size(const gft_ctx* ctx, size_t* bytes);
int gft_checkpoint(gft_ctx* ctx, void* buf, size_t bytes);
int gft_restore(gft_ctx* ctx, const void* buf, size_t bytes);

/* Device-side snapshot for error recovery: gft_snapshot copies the state on
 * the device (allocating one state-sized buffer on its first call);
 * gft_rollback restores the last snapshot and its step counter (*time_iter,
 * may be NULL).  After GFT_EDOMAIN -- where the reference prints "Panic in
 * evolve" and carries on with NaN (WilsonLoops/Source/StandardModel.cpp:474-475)
 * -- rollback returns the lattice to its last valid state. */
int gft_snapshot(gft_ctx* ctx);
int gft_rollback(gft_ctx* ctx, unsigned* time_iter);

/* Evolution.  Replaces Lattice::evolve() (Core/Source/Lattice.cpp:755-876)
 * for nsteps steps.  energy_freq / gauss_freq reproduce the Energy and
 * GaugeCondition cadence (measured on step 0 of the context and then every
 * freq steps; 0 disables).  For the two-slot forms a measured energy step also
 * normalises the last SU(2) product of the magnetic and electric terms, as the
 * reference does (WilsonLoops/Source/StandardModel.cpp:374,433).  rows may be
 * NULL when both frequencies are 0; otherwise it receives one gft_obs per
 * measured step, up to max_rows, and *nrows the count. */
int gft_evolve(gft_ctx* ctx, unsigned nsteps, unsigned energy_freq, unsigned gauss_freq,
               gft_obs* rows, unsigned max_rows, unsigned* nrows);

/* Observables of the current state without stepping (momentum forms: energy at
 * the current time; two-slot forms: the reference's mixed-time energy). */
int gft_observe(gft_ctx* ctx, gft_obs* out);

/* gft_observe without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The observables are those
 * of gft_observe, bitwise: the same kernels, with their per-block reduction
 * partials in a library-static device array of GFT_NOALLOC_SCRATCH_BYTES per
 * device (held until the process ends) instead of a context buffer.  The
 * momentum forms only; lattices of at most GFT_NOALLOC_MAX_SITES sites (more:
 * GFT_EINVAL).  Every call on one device shares that device's array, so calls
 * that overlap on one device (contexts on other streams or threads) race. */
#define GFT_NOALLOC_SCRATCH_BYTES (8u << 20)
#define GFT_NOALLOC_MAX_SITES 22369536LL   /* (8 MiB / (12 doubles per 256-site block)) x 256 */
int gft_observe_noalloc(gft_ctx* ctx, gft_obs* out);

/* Extra lattice observables (FP64): mean 1/2 tr of SU(2) Wilson loops and mean
 * cos of U(1) Wilson loops of size 1x1 and 2x2 over the three spatial planes,
 * and <|phi|^2>. */
typedef struct gft_loops {
    double w11_su2, w11_u1, w22_su2, w22_u1;
    double phi_sq;
} gft_loops;
int gft_wilson_loops(gft_ctx* ctx, gft_loops* out);

/* Launch shape of the momentum-form step kernel: tiles of bz x by sites
 * (bz along z, by along y, bz*by = 256); the block order (0: hardware order,
 * x fastest in (y,z) column tiles; 1: the same dealt to the XCDs in
 * contiguous ranges; 2: each XCD sweeps the x-planes of its own y-band; 3,
 * the default, and 4: as 2 with two or four narrower y-bands per XCD, swept
 * one after the other); and the occupancy target of the FP32 SM kernel (min_waves 3 or 4
 * waves per SIMD, the default being 4; 0 leaves the register allocation
 * unconstrained); and `stage` for the FP32 forms (0: every neighbour read from
 * global memory; 1: each block gathers q, y and phi of its tile and a one-site
 * ring in y and z into LDS, so that only the neighbours in the planes x +- 1
 * are global reads; 2, the default: as 1, and the SU(2) x U(1) plaquettes and
 * the covariant differences (of one or two doublets) inside the x-plane are
 * evaluated once per block and exchanged through LDS instead of by each of the
 * sites that use them, for tiles with bz + by <= 256, otherwise as 1; FP64
 * forms always run unstaged).
 * Every tile shape, order and staging choice gives bitwise the same fields;
 * the autotuner (gftevolver_amd/tune.py) picks the fastest.  No reference
 * counterpart (the reference has no launch configuration). */
int gft_set_launch(gft_ctx* ctx, int bz, int by, int order, int min_waves, int stage);

/* Measurement helper (no reference counterpart): copies `bytes` between two
 * device buffers `reps` times with a plain streaming copy kernel
 * (gft_copy_kernel) and returns the average duration.  Its traffic is known
 * exactly (bytes read + bytes written per launch), so rocprofv3's FETCH_SIZE
 * and WRITE_SIZE of this kernel calibrate the same counters of the step kernel. */
int gft_bench_copy(size_t bytes, int reps, double* avg_ms);

/* Measurement helper (no reference counterpart): the HBM floor of the step's
 * data set.  Streams exactly the arrays one FP32 SM momentum-form step of this
 * context reads and writes (own site only: q, y, phi, P, E and residuals in,
 * the same out), with the step kernel's load and store kinds, tiles and block
 * order (gft_set_launch) and trivial arithmetic, `reps` times, and returns the
 * average launch duration.  It
 * overwrites the context's state: import or generate a state again before
 * evolving (gft_ensemble_evolve refuses a member until then). */
int gft_bench_streams(gft_ctx* ctx, int reps, double* avg_ms);

/* Batched ensembles (no reference counterpart: the reference evolves one
 * lattice per process).  Independent FP32 momentum-form SM lattices of any
 * sizes advance together: one kernel launch per step over a work
 * list of 256-site tiles of all members, optionally replayed from a hipGraph
 * that holds two steps.  Each member is an ordinary context (gft_ensemble_member)
 * for import, export, gft_observe and gft_wilson_loops; its fields are bitwise
 * those of the same member evolved alone.  A member's tile shape comes from
 * gft_ensemble_set_launch or gft_set_launch on the member context (whose
 * `stage` also applies: the ensemble kernel runs the lowest staging mode any
 * member asks for or supports); block order and min_waves do not apply. */
typedef struct gft_ensemble gft_ensemble;
int         gft_ensemble_create(const gft_desc* members, int nmembers, gft_ensemble** out);
void        gft_ensemble_destroy(gft_ensemble* ens);
const char* gft_ensemble_last_error(const gft_ensemble* ens);
int         gft_ensemble_member(gft_ensemble* ens, int member, gft_ctx** ctx);
int         gft_ensemble_set_launch(gft_ensemble* ens, int member, int bz, int by);   /* tile shape per member */
int         gft_ensemble_evolve(gft_ensemble* ens, unsigned nsteps, int use_graph);
int         gft_ensemble_last_step_ms(gft_ensemble* ens, double* avg_ms);

/* Surrogate network (no reference counterpart): an MLP mapping initial-condition
 * and coupling parameters to observable time series, BF16 weights and
 * activations with FP32 accumulation on the MFMA matrix cores, trained on the
 * GPU.  n_in inputs -> n_hidden layers of `width` with GELU -> n_heads linear
 * heads of n_out outputs each, sharing the trunk.  Parameters are FP32 in one
 * flat vector: for each hidden layer l, W_l[width][in_l] (row-major, in_0 =
 * n_in, else width) then b_l[width]; then W_head[n_heads * n_out][width] and
 * b_head[n_heads * n_out].  The loss of a batch is
 *   sum_{s,h,o} w[s][h] (y[s][h][o] - t[s][o])^2 / (n_out * batch * n_heads),
 * with per-head sample weights w (bootstrap resampling of the heads). */
typedef struct gft_mlp_desc {
    int n_in;        /* inputs (8)                                     */
    int width;       /* hidden width, a multiple of 16 (256)            */
    int n_hidden;    /* hidden layers (4)                               */
    int n_heads;     /* output heads (4)                                */
    int n_out;       /* outputs per head, a multiple of 16 (256)        */
    int max_batch;   /* largest batch, a multiple of 16                 */
} gft_mlp_desc;
typedef struct gft_mlp gft_mlp;
int         gft_mlp_create(const gft_mlp_desc* desc, gft_mlp** out);
void        gft_mlp_destroy(gft_mlp* mlp);
const char* gft_mlp_last_error(const gft_mlp* mlp);   /* mlp may be NULL: last create error */
long long   gft_mlp_num_params(const gft_mlp* mlp);
int         gft_mlp_set_params(gft_mlp* mlp, const float* params);   /* also resets the Adam moments */
int         gft_mlp_get_params(gft_mlp* mlp, float* params);
/* y[s][head][o] for x[s][n_in], s < n (any n). */
int         gft_mlp_forward(gft_mlp* mlp, const float* x, int n, float* y);
/* Loss and its gradient (flat, parameter order) for one batch (multiple of 16). */
int         gft_mlp_gradients(gft_mlp* mlp, const float* x, const float* t, const float* w, int batch, float* loss,
                              float* grad);
/* Adam (beta 0.9, 0.999) on n samples, seeded minibatch shuffle per epoch;
 * epoch_loss[e] = mean minibatch loss of epoch e (may be NULL). */
int         gft_mlp_train(gft_mlp* mlp, const float* x, const float* t, const float* w, int n, int batch, int epochs,
                          float lr, unsigned seed, float* epoch_loss);

/* Gaussian-random-field initial state generated on the GPU (no reference
 * counterpart: the reference's initial conditions are RandomUniform and zero,
 * InitialConditions/Source/RandomUniform.cpp:60-77).  Each real component of
 * each doublet is white noise from Philox4x32-10 (counter = (site, component,
 * 0, 0), key = seed; Box-Muller on two 53-bit uniforms) filtered by
 * |k|^(slope/2) for 0 < |k| <= kmax, zero mean and unit rms; each doublet is
 * then (0, 0, vev[k], 0) plus those four fields scaled to a total rms `rms`.
 * The state is at time_iter 0 with both time slots equal (zero scalar momentum),
 * identity links and E = 0.  Every side must be 2^a 3^b and at most 256. */
typedef struct gft_grf_desc {
    unsigned long long seed;
    double rms;        /* sqrt(<sum_c dphi_c^2>) over the four components of a doublet */
    double slope;      /* P(k) ~ |k|^slope                                               */
    double kmax;       /* cut-off; <= 0 selects pi / (2 dx)                              */
    double vev[2];     /* vev of doublet 0 and 1 (component 2)                           */
} gft_grf_desc;
int gft_init_grf(gft_ctx* ctx, const gft_grf_desc* grf);

/* Gaussian initial state generated on the GPU (the 32^3 workload's; no reference
 * counterpart, as for gft_init_grf): each real component c of each doublet is
 * (0, 0, vev[k], 0)[c] + sigma * (standard normal of (site, component)) from
 * the same Philox4x32-10 stream as gft_init_grf's white noise, unfiltered.
 * Time 0, both time slots equal (zero scalar momentum), identity links, E = 0.
 * The reference's own RandomUniform initial condition (std::mt19937, a
 * sequential stream, RandomUniform.cpp:60-77) stays on the host (gft::RandomUniform). */
typedef struct gft_gauss_desc {
    unsigned long long seed;
    double sigma;      /* standard deviation of every real component */
    double vev[2];     /* vev of doublet 0 and 1 (component 2)        */
} gft_gauss_desc;
int gft_init_gaussian(gft_ctx* ctx, const gft_gauss_desc* g);

/* Wait for all queued work on the context's stream. */
int gft_sync(gft_ctx* ctx);

/* Timing: average duration (ms) of a step over the last gft_evolve call, from
 * HIP events on the context's own stream that bracket each run of consecutive
 * step launches (no event between two steps, so the steps run as they would
 * untimed; the average is the step kernel's duration plus the gaps between
 * consecutive steps).  Returns GFT_EINVAL when no step ran. */
int gft_last_step_ms(gft_ctx* ctx, double* avg_ms, double* total_ms, unsigned* nsteps);

#ifdef __cplusplus
}
#endif
#endif /* GFT_H */
