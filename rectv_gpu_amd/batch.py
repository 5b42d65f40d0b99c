"""Many reconstructions on one GPU: the batched form and the serving loop.

A ``Job`` is one call of ``Solver.recon`` (src/rectv_gpu/solver.py:45-84) with
its own sizes: data [nz, ntheta, n], theta [ntheta], phi [m, 2*ntheta], the
same parameters and defaults (lambda0 is multiplied by lambda1, the rotation
centre defaults to n/2) and an optional early stop (``tol``) on the relative
change ||f(k+1) - f(k)|| / ||f(k+1)|| between outer iterations.

``solve_batch`` runs a list of uneven jobs together: solvers are prepared on a
pool of host threads (the host builds a job's geometry tables while earlier
jobs already iterate), every job's outer iteration is captured once as a HIP
graph, and the iterations of all jobs are enqueued a few at a time on their own
streams so that their kernels overlap on the device.  Launch shapes are tuned
(``Solver.autotune``) for jobs with n >= AUTOTUNE_MIN_N; the tuned shapes are
cached per size class, so a batch tunes once per class.  ``batch_iterate`` (the C
entry point ``rtv_batch_iterate``) does the same round robin for solvers a
caller has already prepared.

``Server`` is the serving form: jobs arrive at any time (``submit``), are
admitted while their device footprint fits the memory budget and the number
of active jobs is below ``max_active``, keep their primal/dual state resident
on the device, and have their iterations interleaved with those of the other
active jobs until they reach niter or their tolerance; a finished job frees
its memory for the next admission (continuous batching).  Each job's result
is the same as that of a solve on its own: the same kernels run in the same
order on its own buffers.
"""
import collections
import concurrent.futures as cf
import ctypes
import threading
import time

import numpy as np

from . import _lib


class Job:
    """One reconstruction request (the arguments of Solver.recon plus nzp and tol)."""

    _ids = 0

    def __init__(self, data, theta, phi, rot_center=None, lambda0=1e-3, lambda1=1.0, niter=1, titer=1,
                 nzp=None, tol=None, name=None):
        data = np.ascontiguousarray(data, np.float32)
        if data.ndim != 3:
            raise ValueError("data must be [nz, ntheta, n]")
        self.nz, self.ntheta, self.n = data.shape
        phi = np.ascontiguousarray(phi, np.float32)
        if phi.ndim != 2 or phi.shape[1] != 2 * self.ntheta:
            raise ValueError("phi must be [m, 2*ntheta]")
        self.m = phi.shape[0]
        self.data, self.phi = data, phi
        self.theta = np.ascontiguousarray(theta, np.float32).reshape(self.ntheta)
        self.nzp = self.nz if nzp is None else int(nzp)
        self.rot_center = self.n / 2 if rot_center is None else float(rot_center)
        self.lambda0, self.lambda1 = float(lambda0), float(lambda1)
        self.niter, self.titer = int(niter), int(titer)
        self.tol = tol
        Job._ids += 1
        self.name = name if name is not None else "job%d" % Job._ids
        self.result = None
        self.history = None      # [iterations, 7]: r, s, rho, ||df||^2, ||f||^2, data term, sum |grad f|
        self.iterations = 0
        self.error = None
        self.t_submit = self.t_admit = self.t_start = self.t_done = None
        self.done = threading.Event()

    def fresh(self):
        """A new, unsolved job with the same inputs and parameters."""
        return Job(self.data, self.theta, self.phi, self.rot_center, self.lambda0, self.lambda1, self.niter,
                   self.titer, self.nzp, self.tol, self.name)

    @property
    def voxels(self):
        return self.n * self.n * self.nz * self.m

    def footprint(self):
        from .solver import Solver
        return Solver.footprint(self.n, self.ntheta, self.m, self.nz, self.nzp)

    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


def _converged(job, hist):
    """Relative change of the last finished iteration below job.tol."""
    if job.tol is None or len(hist) == 0:
        return False
    d2, f2 = hist[-1, 3], hist[-1, 4]
    return f2 > 0 and np.sqrt(d2 / f2) < job.tol


def _default_factory(job):
    from .solver import Solver
    return Solver(job.n, job.ntheta, job.m, job.nz, job.nzp, 1)


AUTOTUNE_MIN_N = 512  # autotune="auto": tune jobs of at least this n (small jobs keep the defaults)


def _wants_tuning(job, autotune):
    if autotune == "auto":
        return job.n >= AUTOTUNE_MIN_N
    return bool(autotune)


def prepare(job, factory=_default_factory, graph=True, autotune="auto"):
    """Create the job's solver, upload inputs and geometry, zero the state.
    autotune: True, False or "auto" (only jobs with n >= AUTOTUNE_MIN_N, where the
    tuning launches are small against the job's own work)."""
    s = factory(job)
    try:
        s.set_inputs(job.data, job.theta, job.phi, job.rot_center)
        s.reset()
        if _wants_tuning(job, autotune):
            s.autotune()
            s.reset()
        if graph:
            s.set_graph(True)
    except Exception:
        s.free()
        raise
    return s


def finish(job, s):
    """Fetch result and history of a finished job and free its solver."""
    it = s.iterations()
    job.iterations = it
    job.history = s.history(max(0, it - 1024), min(it, 1024))
    job.result = s.result()
    s.free()


def solve_batch(jobs, workers=8, graph=True, autotune="auto", budget=None, ahead=4):
    """Solve uneven jobs together (see the module docstring): all jobs are
    submitted at once to a Server with a slot for every job, so host
    preparation (geometry tables, uploads) of later jobs overlaps the device
    work of earlier ones and the jobs' iterations are interleaved on their own
    streams.  Jobs whose footprints do not fit ``budget`` (default: 90% of
    free device memory) wait for memory like served jobs.  Returns the jobs."""
    srv = Server(budget=budget, max_active=max(1, len(jobs)), ahead=ahead, workers=workers, graph=graph,
                 autotune=autotune)
    try:
        for j in jobs:
            srv.submit(j)
        srv.run()
    finally:
        srv.close()
    return jobs


def batch_iterate(solvers, niter, titer, lambda0, lambda1):
    """rtv_batch_iterate over prepared solvers (inputs set, state reset):
    niter[j] outer iterations of each, interleaved round robin on their own
    streams; lambda0 is used as given (already multiplied by lambda1)."""
    k = len(solvers)
    hs = (ctypes.c_void_p * k)(*[s._handle().value for s in solvers])
    arr = lambda v, dt: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dt), (k,)))
    ni, ti = arr(niter, np.int32), arr(titer, np.int32)
    l0, l1 = arr(lambda0, np.float32), arr(lambda1, np.float32)
    _lib.check(_lib.lib().rtv_batch_iterate(hs, k, ni.ctypes.data, ti.ctypes.data, l0.ctypes.data, l1.ctypes.data))


def solve_sequential(jobs):
    """The plain form: one blocking Solver.recon per job, one after another."""
    from .solver import Solver
    for j in jobs:
        j.t_submit = j.t_admit = j.t_start = time.perf_counter()
        with Solver(j.n, j.ntheta, j.m, j.nz, j.nzp, 1) as s:
            j.result = s.recon(j.data, j.theta, j.phi, j.rot_center, j.lambda0, j.lambda1, j.niter, j.titer)
        j.iterations = j.niter
        j.t_done = time.perf_counter()
        j.done.set()
    return jobs


def device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise _lib.RectvError("hipMemGetInfo failed")
    return free.value


class Server:
    """Continuous-batching request loop on one GPU (see the module docstring).

    budget: device bytes available to jobs (default 90% of free memory at start);
    max_active: jobs resident at once; ahead: iterations enqueued per job at a
    time; workers: host threads preparing admitted jobs (geometry, upload).
    ``factory`` creates a job's solver (the HIP solver by default) and
    ``footprint`` gives the device bytes a job is charged (Job.footprint).
    """

    def __init__(self, budget=None, max_active=32, ahead=2, workers=4, graph=True, autotune="auto",
                 factory=_default_factory, footprint=None):
        self.budget = budget
        self.max_active = max_active
        self.ahead = max(1, int(ahead))
        self.graph, self.autotune = graph, autotune
        self.factory = factory
        self.footprint = footprint if footprint is not None else (lambda job: job.footprint())
        self._pool = cf.ThreadPoolExecutor(max_workers=workers)
        self._lock = threading.Lock()
        self._queue = collections.deque()
        self._preparing = {}   # future -> (job, bytes)
        self._active = []      # [job, solver, bytes, enqueued, checked]
        self.used = 0
        self.finished = []
        self.peak_active = 0

    def submit(self, job):
        job.t_submit = time.perf_counter()
        job._fp = int(self.footprint(job))
        with self._lock:
            self._queue.append(job)
        return job

    def pending(self):
        with self._lock:
            return len(self._queue) + len(self._preparing) + len(self._active)

    def _admit(self):
        budget = self.budget
        with self._lock:
            while self._queue and len(self._preparing) + len(self._active) < self.max_active:
                # first fit in arrival order
                pick = None
                for j in self._queue:
                    if self.used + j._fp <= budget:
                        pick = j
                        break
                if pick is None:
                    if not self._preparing and not self._active:
                        j = self._queue.popleft()
                        j.error = "job needs %d device bytes, more than the budget %d" % (j._fp, budget)
                        j.t_done = time.perf_counter()
                        self.finished.append(j)
                        j.done.set()
                        continue
                    break
                self._queue.remove(pick)
                pick.t_admit = time.perf_counter()
                self.used += pick._fp
                fut = self._pool.submit(prepare, pick, self.factory, self.graph, self.autotune)
                self._preparing[fut] = (pick, pick._fp)

    def _retire(self, job, nbytes):
        job.t_done = time.perf_counter()
        self.used -= nbytes
        self.finished.append(job)
        job.done.set()

    def step(self):
        """One pass of the loop; returns True if anything progressed."""
        if self.budget is None:
            self.budget = 0.9 * device_free_bytes()
        progressed = False
        self._admit()
        for fut in [f for f in self._preparing if f.done()]:
            job, nb = self._preparing.pop(fut)
            progressed = True
            try:
                s = fut.result()
            except Exception as e:  # a failed preparation fails its job only
                job.error = str(e)
                self._retire(job, nb)
                continue
            job.t_start = time.perf_counter()
            k = min(self.ahead, job.niter)
            try:
                if k > 0:
                    s.iterate_async(k, job.titer, job.lambda0 * job.lambda1, job.lambda1)
            except Exception as e:  # a failing enqueue fails its job only
                self._fail(job, s, nb, e)
                continue
            self._active.append([job, s, nb, k, 0])
        self.peak_active = max(self.peak_active, len(self._active))
        keep = []
        for ent in self._active:
            job, s, nb, enq, chk = ent
            try:
                if not s.idle():
                    keep.append(ent)
                    continue
                progressed = True
                hist = s.history(chk, enq - chk) if enq > chk else np.zeros((0, 7))
                ent[4] = enq
                if enq >= job.niter or _converged(job, hist):
                    finish(job, s)
                    self._retire(job, nb)
                    continue
                k = min(self.ahead, job.niter - enq)
                s.iterate_async(k, job.titer, job.lambda0 * job.lambda1, job.lambda1)
            except Exception as e:  # an error on one job's solver ends that job only
                progressed = True
                self._fail(job, s, nb, e)
                continue
            ent[3] = enq + k
            keep.append(ent)
        self._active = keep
        return progressed

    def _fail(self, job, s, nbytes, err):
        job.error = str(err)
        try:
            s.free()
        except Exception:
            pass
        self._retire(job, nbytes)

    def run(self, until=None, poll=2e-4):
        """Drive the loop until ``until()`` is true (default: nothing pending)."""
        while True:
            if until is None:
                if self.pending() == 0:
                    break
            elif until():
                break
            if not self.step():
                time.sleep(poll)

    def close(self):
        for ent in self._active:
            ent[1].free()
        self._active = []
        for fut, (job, nb) in list(self._preparing.items()):
            try:
                fut.result().free()
            except Exception:
                pass
        self._preparing = {}
        self._pool.shutdown(wait=True)


def latency_stats(jobs):
    lat = np.array([j.latency() for j in jobs if j.latency() is not None])
    if len(lat) == 0:
        return {}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
