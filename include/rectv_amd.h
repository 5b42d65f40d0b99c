/*
 * rectv_amd.h -- C ABI of the MI355X-native rectv solver (librectv_amd.so).
 *
 * Plain C types only: host pointers, sizes, scalars.  Every entry point returns
 * 0 on success and a negative status on failure; rtv_last_error() then returns
 * a message (thread-local).  All host arrays are C-contiguous float32/float64
 * in the reference's layouts:
 *   g      [nz][ntheta][n]          float32   projection data
 *   theta  [ntheta]                 float32   angles (radians)
 *   phi    [m][2*ntheta]            float32   basis, re/im interleaved
 *   fres   [nz][m][n][n]            float32   reconstruction
 * Citations are to the reference nikitinvv/rectv_gpu (paths relative to its root).
 */
#ifndef RECTV_AMD_H
#define RECTV_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtv_solver rtv_solver;

enum rtv_status {
    RTV_OK = 0,
    RTV_ERR_ARG = -1,      /* invalid argument / shape */
    RTV_ERR_HIP = -2,      /* HIP runtime error */
    RTV_ERR_NOMEM = -3,    /* device allocation failed */
    RTV_ERR_STATE = -4     /* call not valid in the current state (e.g. freed) */
};

/* Message of the last failing call on this thread ("" if none). */
const char *rtv_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int rtv_device_count(void);

/* Replaces rectv::rectv(n, ntheta, m, nz, nzp, ngpus), src/cuda/rectv.cu:5-59
 * (SWIG ctor src/cuda/rectv.i:21).  ngpus is clamped like the reference
 * (min(ngpus, nz/nzp), rectv.cu:9) and then to the one GPU this library drives.
 * Fails (instead of truncating, rectv.cu:162) when nz % nzp != 0. */
int rtv_create(int n, int ntheta, int m, int nz, int nzp, int ngpus, rtv_solver **out);

/* Replaces rectv::free(), src/cuda/rectv.cu:63-91 (idempotent; never resets the
 * device).  Also releases the handle itself. */
int rtv_free(rtv_solver *h);

/* Replaces rectv::run(fres, g, theta, phi, center, lambda0, lambda1, niter, titer),
 * src/cuda/rectv.cu:102-252.  lambda0 is used as given (the Python layer
 * multiplies it by lambda1 first, src/rectv_gpu/solver.py:76).  Prints the
 * reference's per-iteration line to stderr unless RECTV_QUIET=1. */
int rtv_run(rtv_solver *h, float *fres, const float *g, const float *theta, const float *phi,
            float center, float lambda0, float lambda1, int niter, int titer);

/* Replaces rectv::check_adjoints(res, g, theta, phi, lambda1, center),
 * src/cuda/rectv.cu:255-289 (requires nz == nzp, as the reference states at :258).
 * res receives 6 doubles. */
int rtv_check_adjoints(rtv_solver *h, double *res, const float *g, const float *theta,
                       const float *phi, float lambda1, float center);

/* ---- staged form of rtv_run (same arithmetic), used to time the solver with its
 * inputs already resident in device memory.  rtv_run == set_inputs + reset +
 * iterate + get_result.  No reference counterpart: the reference fuses these
 * steps inside rectv::run (rectv.cu:107-125 upload/zero, :140-246 loop, :250-251
 * download). */
int rtv_set_inputs(rtv_solver *h, const float *g, const float *theta, const float *phi, float center);
int rtv_reset(rtv_solver *h);
/* Runs niter outer iterations.  hist (may be NULL) receives niter rows of
 * {r, s, rho used, rdiff = ||f(k+1)-f(k)||^2}. */
int rtv_iterate(rtv_solver *h, int niter, int titer, float lambda0, float lambda1, double *hist);
int rtv_get_result(rtv_solver *h, float *fres);
int rtv_synchronize(rtv_solver *h);

/* ---- solver state (checkpoint / restore).  No reference counterpart: rectv::run
 * always starts from zero (src/cuda/rectv.cu:119-125).  Layouts: f [nz][m][n][n]
 * float32 (the previous iterate, from which halo slabs are read);
 * psi, mu [nz/nzp][nzp+1][m+1][n+1][n+1][4] float32 (the reference's float4 dual
 * arrays per chunk, components x, y, t, z, ghost entries included);
 * dchi = div(mu - rho*psi) [nz][m][n][n]. */
/* Sets the state (NULL f / psi / mu: zero), forms div(mu - rho psi) for lambda1 and
 * restarts the iteration counter and history; rtv_iterate then continues from it. */
int rtv_set_state(rtv_solver *h, const float *f, const float *psi, const float *mu, float rho, float lambda1);
/* Copies the state out (any pointer may be NULL; rho: 1 float).  The solver keeps
 * chi = mu - rho psi in place of psi (the dual part of the inner steps), so psi is
 * returned as (mu - chi) / rho, equal to the psi of the last closing phase to FP32
 * rounding; a set_state / get_state round trip returns f and mu bitwise. */
int rtv_get_state(rtv_solver *h, float *f, float *psi, float *mu, float *dchi, float *rho);

/* ---- filtered back-projection and support mask: the reference's unused USFFT
 * extras applyfilter (src/include/kernels_radonusfft.cuh:173-189), the `filter`
 * argument of radonusfft::adjR (src/cuda/radonusfft.cu:106) and circ
 * (kernels_radonusfft.cuh:146-159). */
/* r >= 0: keep pixels with 4x^2 + 4y^2 < 1 - r, x = (col - n/2)/n (circ), in the
 * filtered back-projection and after every inner step (a support constraint);
 * r < 0 (the default): no mask, the reference's iteration. */
int rtv_set_support(rtv_solver *h, float r);
/* f [nz][m][n][n] = in every frame pi/(4 sqrt(m)) Re adjR_w(g): the filtered
 * back-projection of each slice of g [nz][ntheta][n] as a static object (geometry
 * from rtv_set_inputs), masked by the support when set. */
int rtv_apply_fbp(rtv_solver *h, const float *g, float *f);
/* Starting state from the filtered back-projection of the resident data (instead of
 * f = 0, src/cuda/rectv.cu:119-125): f = FBP, psi = mu = 0, rho = 0.5; then
 * rtv_iterate continues from it. */
int rtv_fbp_start(rtv_solver *h);

/* ---- single steps of the ADMM iteration on the resident state (parity tests of
 * the fused kernels; src/cuda/admm.cu:74-122 split at its two phases). */
/* One inner step (admm.cu:85-97) of every chunk applied to fcur [nz][m][n][n]:
 * fout = fcur - 1/2 div(rho(grad fcur - psi) + mu) - 1/2 R_apr^*(R_apr fcur - g),
 * halo slabs from the resident previous iterate; rtv_set_inputs supplies g. */
int rtv_apply_inner_step(rtv_solver *h, const float *fcur, float *fout, float lambda1);
/* The closing phase (admm.cu:101-117 with the rho rule of rectv.cu:224-232) applied
 * to fcur as the new iterate: updates the resident psi, mu, rho and
 * div(mu - rho psi); row (may be NULL) receives the history row (7 doubles, see
 * rtv_history).  The resident previous iterate is unchanged. */
int rtv_apply_prox(rtv_solver *h, const float *fcur, float lambda0, float lambda1, double *row);

/* rtv_apply_prox on caller-owned DEVICE arrays, allocating nothing: no reference
 * counterpart; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as rtv_apply_prox.  fcur, fprev [nz][m][n][n]; psi, mu
 * [nz/nzp][nzp+1][m+1][n+1][n+1][4] (16-byte aligned, updated in place); dchi
 * [nz][m][n][n] out; rho: 1 float in/out; row: 7 doubles out (as rtv_history, data
 * term NaN).  Work is enqueued on `stream` (a hipStream_t; NULL = null stream) and
 * the call returns without waiting.  Its scratch (partial sums) is a fixed 8 MiB
 * array per device, held until the process ends: sizes needing more,
 * 40 * ceil((n+1)/16)(m+1)(nzp+1)(nz/nzp) + 72 bytes, are rejected (RTV_ERR_ARG).
 * Every call on one GPU shares that array, so calls that overlap in time on one
 * GPU (other streams, other threads) race and must be serialised by the caller. */
int rtv_prox_noalloc(const float *fcur, const float *fprev, float *psi, float *mu, float *dchi, float *rho,
                     double *row, int n, int m, int nz, int nzp, float lambda0, float lambda1, void *stream);

/* ---- operator entry points (parity tests).  Each applies one reference operator to
 * all nz slices, with inputs/outputs in host memory; rtv_set_inputs must have
 * been called (it supplies theta, phi and center). */
/* rectv::radonapr(g, f, tau=1), src/cuda/operators.cu:6-32:  g = R_apr f. */
int rtv_apply_radon(rtv_solver *h, const float *f, float *g);
/* rectv::radonapradj(f, g, tau=1) into zero, src/cuda/operators.cu:34-58:  f = R_apr^* g. */
int rtv_apply_radon_adj(rtv_solver *h, const float *g, float *f);
/* rectv::gradient(h2, f, lambda1, iz), src/cuda/operators.cu:60-66 for chunk iz:
 * fext [nzp+2][m][n][n] (slab 0 / nzp+1 = halos, read when the chunk has a
 * neighbour); h2 out [nzp+1][m+1][n+1][n+1][4] (the reference's float4 order). */
int rtv_apply_gradient(rtv_solver *h, const float *fext, int iz, float lambda1, float *h2);
/* rectv::divergent(fn, h2, lambda1, tau), src/cuda/operators.cu:68-71:
 * f [nzp][m][n][n] += tau * div(h2). */
int rtv_apply_divergence(rtv_solver *h, const float *h2, float lambda1, float tau, float *f);

/* ---- profiling: per-kernel HIP-event timing on the solver's stream. */
int rtv_set_profiling(rtv_solver *h, int on);
/* Writes "name count total_ms\n" lines into buf (NUL-terminated, truncated to len). */
int rtv_get_profile(rtv_solver *h, char *buf, int len);
/* Device bytes held by the solver (persistent state + scratch). */
long long rtv_device_bytes(rtv_solver *h);

/* ---- asynchronous, graph-replayed and batched forms (serving / batching).  No
 * reference counterpart: the reference runs one blocking solve per process
 * (rectv::run, rectv.cu:102-252) and keeps rho on the host (rectv.cu:125, 220-232);
 * here rho, its rule and the per-iteration norms stay on the device, so outer
 * iterations can be enqueued without host round trips. */
/* Enqueue niter outer iterations on the solver's stream and return at once. */
int rtv_iterate_async(rtv_solver *h, int niter, int titer, float lambda0, float lambda1);
/* *idle = 1 when all enqueued work of the solver has finished, else 0. */
int rtv_query(rtv_solver *h, int *idle);
/* Outer iterations enqueued since the last rtv_reset. */
int rtv_iterations(rtv_solver *h, long long *count);
/* Rows [first, first+count) since reset (waits for the solver's stream), 7 doubles
 * each: r, s, rho used, ||f(k+1)-f(k)||^2, ||f(k+1)||^2, 1/2||R_apr f(k) - g||^2 (the
 * data term of the iterate entering iteration k; NaN when titer = 0), and
 * sum |grad f(k+1)| over the dual positions of each chunk (halos from f(k)).  The
 * objective 1/2||R f - g||^2 + lambda0/2 sum|grad f| of iterate k is column 5 of
 * row k plus lambda0/2 times column 6 of row k-1.  The last 1024 rows are held. */
int rtv_history(rtv_solver *h, long long first, int count, double *rows);
/* on != 0: each outer iteration is captured once as a HIP graph and replayed. */
int rtv_set_graph(rtv_solver *h, int on);
/* Times the launch-shape candidates of the gather and scatter on this solver's
 * geometry and keeps the fastest (cached in the process per n, M and power-of-two
 * class of ntheta).  Results are bitwise independent of the choice.  Outputs (may
 * be NULL): slices per workgroup. */
int rtv_autotune(rtv_solver *h, int *gather_slices, int *scatter_slices);
/* Runs niter[j] outer iterations of every solver hs[j] (distinct, inputs set),
 * interleaved round robin on their own streams; returns when all have finished. */
int rtv_batch_iterate(rtv_solver **hs, int count, const int *niter, const int *titer,
                      const float *lambda0, const float *lambda1);
/* Device bytes a solver of these sizes holds (upper bound), for admission control. */
int rtv_footprint(int n, int ntheta, int m, int nz, int nzp, long long *bytes);

/* ---- pre-processing of projection data before the solver (on the GPU, host arrays in
 * and out; no solver handle).  The reference leaves these steps to its examples and
 * to tomocupy: examples/tao/tao.py:34-43, 58-71 and examples/tao/rec_line.txt:1. */
/* out [nz][ntheta][n] = -log((data - mean(dark)) / (1e-3 + mean(flat) - mean(dark)))
 * (take_log = 0: the ratio itself) from data [ntheta][nz][n], dark [ndark][nz][n],
 * flat [nflat][nz][n], frame means over the first axis (tao.py:34-43, including its
 * swapaxes(0, 1) to sinograms). */
int rtv_prep_darkflat(const float *data, const float *dark, const float *flat, int ntheta, int nz, int n,
                      int ndark, int nflat, int take_log, float *out);
/* In place on sinograms g [nz][ntheta][n] (n even, at most 4096): the stripe removal of
 * tao.py:58-71 with its beta and integer mask_size (columns [n/2 - ms/2, n/2 + ms/2),
 * ms = mask_size * n, under Python's slice rules). */
int rtv_prep_stripe_ti(float *g, int nz, int ntheta, int n, float beta, int mask_size);
/* In place on projections proj [ntheta][ny][nz]: single-material phase retrieval
 * (Paganin et al. 2002) as tomopy's retrieve_phase with pad=True applies it:
 * edge-extended to power-of-two-type sizes, multiplied in 2D Fourier space by the
 * normalised filter 1/(lambda dist |k|^2/(4 pi) + alpha), cropped.  pixel_size and
 * dist in cm, energy in keV; padded sizes must factor into 2, 3, 5, 7 (<= 4096). */
int rtv_prep_paganin(float *proj, int ntheta, int ny, int nz, float pixel_size, float dist, float energy,
                     float alpha);

#ifdef __cplusplus
}
#endif
#endif /* RECTV_AMD_H */
