/*
 * speckle_hip.h -- C ABI of the MI355X (gfx950) iterative phase-retrieval
 * library.  Plain pointers and sizes only; every device pointer is a HIP
 * device allocation owned by the caller (e.g. a torch ROCm tensor's
 * data_ptr()), every host pointer is ordinary host memory.
 *
 * The reference has no C ABI: its operator boundary is the name-based kernel
 * dispatch gpu.common._kexec (speckle/gpu.py:295-372) plus the third-party
 * FFT plan pyfft.cl.Plan (built at speckle/phasing.py:372-374, executed at
 * speckle/phasing.py:1074-1075).  Each entry point below names the reference
 * interface it replaces.
 *
 * Return codes: 0 on success, a negative SPK_E* value on failure; the message
 * of the last failure on the calling thread is spk_last_error().  No entry
 * point exits or aborts the process (the reference calls exit() on fatal
 * errors, speckle/gpu.py:366-370, speckle/phasing.py:252-254).
 *
 * Threading: a context owns one HIP stream; all work of its plans is queued
 * on that stream in call order and is asynchronous unless stated.  Different
 * contexts may be driven from different host threads.
 */
#ifndef SPECKLE_HIP_H
#define SPECKLE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPK_OK = 0,
  SPK_EINVAL = -1,   /* bad argument (shape, size, pointer, enum)          */
  SPK_EHIP = -2,     /* a HIP runtime call failed                            */
  SPK_ENOMEM = -3,   /* device allocation failed                             */
  SPK_ESTATE = -4,   /* call out of order (e.g. iterate before set_modulus)  */
  SPK_EUNSUP = -5    /* size or option not built into this library           */
};

/* algorithms of phasing._iteration (speckle/phasing.py:1329-1331) */
enum { SPK_HIO = 0, SPK_ER = 1, SPK_RAAR = 2, SPK_SF = 3, SPK_DM = 4,
       SPK_NOOP = 15 /* frame idle for this step (serving) */ };

/* ER / SF formulas: the reference's numpy branch (phasing.py:1256-1258,
 * 1281-1284) or its OpenCL kernels (phasing_er.cl:13-15, phasing_sf.cl:16-17) */
enum { SPK_SEM_NUMPY = 0, SPK_SEM_OPENCL = 1 };

typedef struct spk_ctx spk_ctx;
typedef struct spk_plan spk_plan;

/* One iteration of one frame: algorithm, feedback beta, semantics, and an
 * optional in-loop shrinkwrap after the update (sigma > 0 enables it). */
typedef struct spk_step {
  int32_t alg;        /* SPK_HIO ... SPK_DM, SPK_NOOP                      */
  int32_t semantics;  /* SPK_SEM_NUMPY or SPK_SEM_OPENCL                    */
  float beta;         /* iterate(beta=) of phasing.py:181                    */
  float sw_sigma;     /* > 0: shrinkwrap after this iteration, Gaussian sigma */
  float sw_tau;       /* shrinkwrap threshold as a fraction of the max blur  */
  int32_t pad_;
} spk_step;

/* Kernel shape choices for a plan (autotuner output).  Zero = library default. */
typedef struct spk_tune {
  int32_t tile;       /* tile edge of the internal layout: 2, 4 or 8         */
  int32_t ept;        /* values per thread of one transform: 8 or 16         */
  int32_t persist;    /* 1: run iteration blocks as one persistent launch    */
  int32_t sw_separate;/* 1: run each shrinkwrap as its own kernels (default:
                         fused into the passes of its step and the next
                         where the blur radius is <= 24 px, i.e. sigma <= 4) */
  int32_t resident;   /* one frame of 256^2..1024^2 in tile 4 (256^2 and
                         512^2 also in tile 2), 8 values per
                         thread: runs of non-DM steps run as one cooperative
                         resident launch each (the frame's modulus, support
                         and iterate held on chip, Phi exchanged through a
                         grid barrier, shrinkwraps inside the launch); 0 or
                         1: where built (the default), 2: never.  A resident
                         launch needs the whole device: no other cooperative
                         launch may run beside it (spk_plan_check reports a
                         barrier that gave up waiting)                       */
  int32_t exchange;   /* resident launches: how Phi crosses the barrier,
                         1 plain + device-scope fences, 2 device-scope
                         accesses + fences, 3 device-scope accesses without
                         fences (the default); 0 = default                   */
  int32_t barrier_polls; /* resident and persistent launches: poll rounds of
                         a grid-barrier wait before it gives up (the launch
                         then drains without waiting and spk_plan_check
                         reports it); 0 = default, 262144.  Small values
                         exist to exercise the give-up path.                 */
  int32_t reserved;
} spk_tune;

/* -------------------------------------------------------------------------
 * Context and device.  Replaces gpu.init / gpu.common.start
 * (speckle/gpu.py:399-421, 491-558): binds one device and one stream.
 * stream: a hipStream_t (e.g. torch.cuda.current_stream().cuda_stream), or
 * NULL for the device's null stream.  The context does not own the stream.
 * ---------------------------------------------------------------------- */
int spk_ctx_create(int device, void* stream, spk_ctx** out);
/* releases the caller's reference; plans hold their own, so the context is
 * freed when it and all its plans have been destroyed, in any order */
int spk_ctx_destroy(spk_ctx* ctx);
int spk_ctx_set_stream(spk_ctx* ctx, void* stream);
/* blocks until all work queued on the context's stream has finished */
int spk_sync(spk_ctx* ctx);
/* message of the last failed call on this host thread ("" if none) */
const char* spk_last_error(void);
/* library version string and the sizes built in (bitmask of log2 N) */
const char* spk_version(void);
uint32_t spk_supported_sizes(void);

/* The layout tile, values per thread and launch mode a plan was built with
 * (the defaults resolved).  No reference counterpart. */
int spk_plan_info(const spk_plan* plan, int* tile, int* ept, int* persist);
/* Shrinkwraps the plan's spk_iterate calls have run so far, fused into the
 * passes and as stand-alone kernels.  No reference counterpart. */
int spk_plan_sw_counts(const spk_plan* plan, long long* fused, long long* separate);

/* Waits for the plan's queued work and reports whether a grid barrier of a
 * resident launch or a persistent iteration block gave up waiting (SPK_EHIP;
 * the flag is then cleared).  No counterpart in the reference, whose
 * kernels each end with a blocking wait (gpu.py:349). */
int spk_plan_check(spk_plan* plan);

/* -------------------------------------------------------------------------
 * Plans.  A plan holds `batch` frames of n x n (n a power of two, as pyfft
 * requires, speckle/phasing.py:1469) resident on the device: iterate psi,
 * its row transform, modulus, support, twiddles and workspace.  Replaces the
 * buffers allocated by phasing.load (speckle/phasing.py:360-374) and the
 * pyfft Plan made there.  shared != 0: one modulus and one support serve
 * every frame (independent trials of one pattern); otherwise each frame has
 * its own (serving slots).
 * ---------------------------------------------------------------------- */
int spk_plan_create(spk_ctx* ctx, int n, int batch, int shared,
                    const spk_tune* tune, spk_plan** out);
int spk_plan_destroy(spk_plan* plan);
/* Bytes of device memory a plan holds once the options it will use have
 * allocated their buffers (for HBM-budget admission).  options: bitwise OR of
 * SPK_PB_* (1 keeps the meaning of the former with_dm argument). */
enum {
  SPK_PB_DM = 1,         /* DM steps: the modified estimate phib (with an
                            ipsf or spectrum also the work buffers of DM's
                            second projection)                              */
  SPK_PB_SHRINKWRAP = 2, /* in-loop shrinkwrap: blur work maps, band maxima */
  SPK_PB_IPSF = 4,       /* spk_set_ipsf: the ipsf and per-frame work frames */
  SPK_PB_SPECTRUM = 8,   /* spk_set_spectrum: per-frame intensity maps       */
  SPK_PB_TRACE = 16      /* error trace of spk_iterate calls of up to 64
                            steps: the per-band FP64 partials              */
};
int spk_plan_bytes(int n, int batch, int shared, int options, size_t* bytes);

/* -------------------------------------------------------------------------
 * Data in and out of a plan.  All array arguments are device pointers to
 * row-major n x n arrays.  frame = -1 addresses every frame (for a shared
 * plan the single modulus / support).
 * ---------------------------------------------------------------------- */
/* phasing.load(modulus=) data path (phasing.py:347-376); also computes the
 * FP64 sum of m^2 used by the modulus error */
int spk_set_modulus(spk_plan* plan, int frame, const float* modulus);
/* phasing.load(support=) data path (phasing.py:389-398); the caller has
 * normalised the support to max 1 as the reference does */
int spk_set_support(spk_plan* plan, int frame, const float* support);
int spk_get_support(spk_plan* plan, int frame, float* support);
/* phasing.seed data path (phasing.py:1026-1031): psi = guess, times the
 * support when mult_support != 0 (common_multiply_f_f2.cl:10-13) */
int spk_set_psi(spk_plan* plan, int frame, const float* guess_c64,
                int mult_support);
/* phasing.get(psi_in) (gpu.py:389-397) */
int spk_get_psi(spk_plan* plan, int frame, float* psi_c64);

/* Transverse partial coherence (phasing.py:414-425 load(ipsf=), divisor at
 * phasing.py:1175-1183): ipsf is an n x n c64 array (row-major, the transform
 * of the point-spread function, corner-centred), shared by the plan's
 * frames.  While set, every iteration's modulus replacement divides by
 * div = sqrt(|ifft2(fft2(|Psi|^2) * ipsf)|) instead of |Psi| and the error
 * trace compares div with m.  NULL switches it off.  DM steps project their
 * modified estimate with its own partially coherent divisor. */
int spk_set_ipsf(spk_plan* plan, const float* ipsf_c64);

/* Longitudinal (spectral) partial coherence (phasing.py:430-448
 * load(spectrum=), applied at phasing.py:1140-1156 with
 * phasing_radial_rescale.cl): K rescale factors (energy / energy of the peak
 * weight) and K weights, HOST arrays, copied.  While set, the squared
 * divisor (|Psi|^2, or the ipsf-blurred intensity when an ipsf is also set)
 * is replaced by sum_k w_k * (its bilinear interpolation at coordinates /
 * r_k) before the square root.  rescales = NULL or K = 0 switches it off.
 * DM steps project their modified estimate with its own rescaled divisor. */
int spk_set_spectrum(spk_plan* plan, const float* rescales, const float* weights, int K);

/* -------------------------------------------------------------------------
 * The loop.  Replaces phasing.iterate's per-iteration work
 * (speckle/phasing.py:260-268 -> _iteration, 1088-1353: pyfft forward,
 * abs_f2_f2, fourier, pyfft inverse, and hio/er/raar/sf/dm1/dm2 kernels).
 *
 * steps: host array.  per_frame == 0: nsteps entries applied to every frame;
 * per_frame != 0: nsteps * batch entries, steps[k * batch + f] for frame f.
 * err_trace: device pointer to nsteps * batch doubles (may be NULL):
 * err_trace[k * batch + f] = sqrt(sum(|F psi_k| - m)^2 / sum m^2) for the
 * iterate psi_k entering step k, accumulated in FP64.
 * ---------------------------------------------------------------------- */
int spk_iterate(spk_plan* plan, const spk_step* steps, int nsteps,
                int per_frame, double* err_trace);

/* -------------------------------------------------------------------------
 * Frames of different sizes in one launch.  spk_group_iterate runs, for each
 * of `count` plans (at most 16, one context), the spk_iterate call
 * (steps[m], nsteps, per_frame[m], err_trace[m]; per_frame and err_trace
 * may be NULL), with the single-projection passes of every plan merged per
 * step into one column-pass and one row-pass launch over all their active
 * frames.  Results equal each plan's own spk_iterate call bitwise.  Steps
 * with DM work or a partially coherent projection run with the plan's own
 * launches in the same call; shrinkwraps as in spk_iterate.  Member plans
 * hold frames of 256^2, 512^2 or 1024^2 in the shape spk_group_shape gives
 * (the layout tile that makes every size's workgroup 256 threads), are
 * not persistent and are not being captured.
 * Pattern: the reference runs one frame size per phasing object and its
 * trials one after another (phasing.py:353, examples/basic_phasing_example.py:27-30);
 * no reference counterpart.
 * ---------------------------------------------------------------------- */
typedef struct spk_group spk_group;
int spk_group_shape(int n, int* tile, int* ept);
int spk_group_create(spk_ctx* ctx, spk_group** out);
int spk_group_destroy(spk_group* group);
int spk_group_iterate(spk_group* group, spk_plan* const* plans, int count,
                      const spk_step* const* steps, int nsteps, const int* per_frame,
                      double* const* err_trace);
/* Merged launches made so far and the member passes they covered. */
int spk_group_counts(const spk_group* group, long long* launches, long long* member_passes);

/* Shrinkwrap now, on every frame: support *= [G_sigma (*) |psi| > tau*max].
 * Build-defined in-loop form of refine_support (phasing.py:1747-1815). */
int spk_shrinkwrap(spk_plan* plan, int frame, float sigma, float tau);

/* ROI copy into a savebuffer slot, phasing.iterate (phasing.py:279-287) ->
 * phasing_copy_to_buffer_f2.cl:6-19.  dst: device c64 (slots, rows, cols),
 * row-major; writes slot `slot`.  Bounds are checked (the kernel of the
 * reference has none). */
int spk_copy_roi(spk_plan* plan, int frame, float* dst_c64, int slots,
                 int slot, int r0, int c0, int rows, int cols);

/* -------------------------------------------------------------------------
 * HIP graphs.  spk_graph_begin starts stream capture on the plan's stream
 * (which must not be the null stream); the plan calls made until
 * spk_graph_end (spk_set_support / spk_set_psi / spk_iterate with
 * per_frame = 0 / spk_copy_roi) are recorded instead of run, and
 * spk_graph_end instantiates them.  spk_graph_launch replays the recorded
 * work on the plan's stream with the same pointers and step coefficients.
 * Workspace must already exist: run the same calls once before capturing.
 * The step tables of the recorded spk_iterate calls (at most 4 MiB per
 * graph) live in buffers the graph owns, so eager calls between replays do
 * not change what a replay runs; spk_graph_destroy frees them.
 * No counterpart in the reference (its launches block one by one,
 * gpu.py:349).
 * ---------------------------------------------------------------------- */
typedef struct spk_graph spk_graph;
int spk_graph_begin(spk_plan* plan);
int spk_graph_end(spk_plan* plan, spk_graph** out);
int spk_graph_launch(spk_graph* graph);
int spk_graph_destroy(spk_graph* graph);

/* -------------------------------------------------------------------------
 * Standalone transform.  Replaces pyfft.cl.Plan.execute(data_in, data_out,
 * inverse) (phasing.py:1074-1075): count frames of n x n c64, row-major,
 * out-of-place or in-place.  Forward unnormalised; inverse scaled by 1/n^2
 * (numpy's ifft2 convention, phasing.py:1078-1083).
 * ---------------------------------------------------------------------- */
int spk_fft2(spk_plan* plan, const float* in_c64, float* out_c64, int count,
             int inverse);

/* -------------------------------------------------------------------------
 * Post-processing of trial stacks (module functions of speckle.phasing).
 * data: device c64 (frames, rows, cols), row-major, modified in place.
 * ---------------------------------------------------------------------- */
/* align_global_phase (phasing.py:1355-1404), closed form, FP64 sums */
int spk_align_global_phase(spk_ctx* ctx, float* data_c64, int frames,
                           int rows, int cols);

/* spk_align_global_phase without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results as spk_align_global_phase; the partial sums live in
 * a static per-device workspace of 2 MiB instead of the context's grown
 * buffer, so frames x 5 x min(64, ceil(rows*cols/256)) must not exceed
 * 262144 (e.g. 819 frames of 128 x 128 or more; SPK_EINVAL beyond).  Calls
 * that overlap on one GPU share the workspace and race. */
int spk_align_global_phase_noalloc(spk_ctx* ctx, float* data_c64, int frames,
                                   int rows, int cols);

/* refine_support's blur (phasing.py:1787-1790): out = Re ifft2(fft2(K) *
 * fft2(in)) for K = fftshift(shape.gaussian((rows, cols), (sigma, sigma)))
 * normalised to sum 1 (shape.py:410-530), computed as the equivalent
 * periodic separable convolution.  in/out: device float (rows, cols). */
int spk_blur2d(spk_ctx* ctx, const float* in, float* out, int rows, int cols,
               float sigma);

/* PRTF (phasing.py:1406-1461 / gpu_prtf 1463-1519): stack (frames, rows,
 * cols) c64 zero-padded top-left into the plan's n x n, transformed, and
 * out = fftshift(|mean_f F_f / |F_f||) as the reference returns it (no guard
 * where |F| = 0, as in the reference).  out: n x n float. */
int spk_prtf(spk_plan* plan, const float* stack_c64, int frames, int rows,
             int cols, float* out);

/* RFTF (phasing.py:1521-1609 without hot-pixel removal or q-average):
 * estimate (rows, cols) c64 zero-padded into n x n, |F| compared with the
 * goal modulus at both alignments, out = sqrt(1/(1+err)) centre-shifted as
 * the reference returns it; scale != 0 maps [1/sqrt(2), 1] to [0, 1].
 * ipsf (n x n c64, or NULL): |F| is first replaced by
 * sqrt(|ifft2(fft2(|F|) * ipsf)|) (phasing.py:1567-1573). */
int spk_rftf(spk_plan* plan, const float* estimate_c64, int rows, int cols,
             const float* goal, const float* ipsf_c64, int scale, float* out);

/* gpu_rftf (phasing.py:1611-1708, phasing_rftf_accumulate.cl:15-24): the
 * per-frame value 1/sqrt(1 + (g - |F_f|)^2 / g^2) with the goal at the corner,
 * averaged over the stack (frames, rows, cols) c64 zero-padded into n x n;
 * out n x n float, centre-shifted as the reference returns it.  ipsf (or
 * NULL): |F_f| is first replaced by sqrt(|ifft2(fft2(|F_f|^2) * ipsf)|)
 * (phasing.py:1676-1690). */
int spk_rftf_stack(spk_plan* plan, const float* stack_c64, int frames, int rows,
                   int cols, const float* goal, const float* ipsf_c64, float* out);

/* Coherence estimation, optimize_coherence (phasing.py:507-930).
 * Richardson-Lucy (method 'richardson_lucy', phasing.py:700-854): one
 * estimate frame (rows, cols) c64 embedded top-left, the n x n modulus;
 * `iterations` updates u *= ifft2(fft2(d / ifft2(fft2(u) P)) Ph) from the
 * start Gaussian; ipsf = fft2(u) (real part with SPK_SEM_NUMPY, the CPU
 * branch) / |ipsf(0,0)|, n x n c64 row-major, not centre-shifted.
 * is_fourier: the estimate is a Fourier-space array (the GPU branch). */
int spk_coherence_rl(spk_plan* plan, const float* estimate_c64, int rows, int cols,
                     const float* modulus, int iterations, int semantics, int is_fourier,
                     float* ipsf_c64);
/* Gaussian fit (method 'gaussian', phasing.py:657-698, 785-803): prepare the
 * autocorrelation of one estimate frame, then evaluate the objective
 * sum (sqrt|fft2(g * ac)| - m)^2 for coherence lengths (cl0, cl1) in pixels
 * (value: a HOST double; the call waits for the result). */
int spk_coherence_gauss_prepare(spk_plan* plan, const float* estimate_c64, int rows, int cols,
                                const float* modulus);
int spk_coherence_gauss_eval(spk_plan* plan, double cl0, double cl1, double* value);

/* center_of_mass_average (phasing.py:1710-1745): out (rows, cols) c64 = the
 * SUM over frames of each frame rolled by minus its integer centre-of-mass
 * offset (of |frame|, truncated) from frame 0's.  Waits for the result. */
int spk_com_average(spk_ctx* ctx, const float* data, int frames, int rows, int cols,
                    float* out);

/* trial average: out (rows, cols) c64 = mean over frames of data; the
 * reference averages aligned trials with np.mean / np.sum on the host
 * (phasing.py:960-961, basic_phasing_example.py:43); FP64 sums */
int spk_stack_mean(spk_ctx* ctx, const float* data_c64, int frames, int rows,
                   int cols, float* out_c64);

/* -------------------------------------------------------------------------
 * Measurement helpers.
 * ---------------------------------------------------------------------- */
/* plain device copy of `bytes` (multiple of 16), for HBM counter calibration */
int spk_copy_bytes(spk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Per-pass timing inside spk_iterate: with enable != 0 every column-pass and
 * row-pass launch of this plan is made with hipExtLaunchKernelGGL and a
 * start/stop HIP event pair recorded by the dispatch itself (the kernel's own
 * execution interval); resident launches get an event pair around their
 * launch on the stream.  spk_plan_profile_read synchronises, returns the
 * number of launches and their summed milliseconds for pass 0 (column), 1
 * (row) or 2 (resident launches), and clears that record. */
int spk_plan_profile(spk_plan* plan, int enable);
int spk_plan_profile_read(spk_plan* plan, int pass, int* launches,
                          double* total_ms);

/* The plan's resident launches (spk_tune.resident) so far: launches and the
 * steps they ran.  No reference counterpart. */
int spk_plan_resident_counts(const spk_plan* plan, long long* launches, long long* steps);

/* Launch `reps` column passes (which = 0), row passes (1) or both (2) of a
 * single-projection step on the plan's current state, timed with HIP events
 * on the context's stream; *ms_out = elapsed milliseconds.  Advances the
 * plan's state like real iterations.  Used to time kernels. */
int spk_time_passes(spk_plan* plan, int which, int reps, const spk_step* step,
                    float* ms_out);

#ifdef __cplusplus
}
#endif

#endif /* SPECKLE_HIP_H */
