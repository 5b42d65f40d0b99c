"""Host logic of the serving loop (no GPU): admission by memory footprint and
slot count, continuous admission as jobs finish, early stop on the relative
change, failure isolation.  The solver is replaced by a stand-in with the same
method set whose iterations complete after a fixed delay."""
import threading
import time

import numpy as np

from rectv_gpu_amd.batch import Job, Server


class FakeSolver:
    """Stand-in for rectv_gpu_amd.Solver: each enqueued iteration takes `dt`
    seconds; the relative change of iteration k is 0.5**k."""

    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, job, dt=2e-3, fail=False):
        self.job, self.dt, self.fail = job, dt, fail
        self.enq = 0
        self.ready_at = 0.0
        with FakeSolver.lock:
            FakeSolver.live += 1
            FakeSolver.peak = max(FakeSolver.peak, FakeSolver.live)
        self.freed = False

    def set_inputs(self, data, theta, phi, c):
        if self.fail:
            raise ValueError("bad inputs")

    def reset(self):
        self.enq = 0

    def autotune(self):
        return 8, 4

    def set_graph(self, on=True):
        pass

    def iterate_async(self, k, titer, l0, l1):
        self.ready_at = max(self.ready_at, time.perf_counter()) + k * self.dt
        self.enq += k

    def idle(self):
        return time.perf_counter() >= self.ready_at

    def iterations(self):
        return self.enq

    def history(self, first, count):
        k = np.arange(first + 1, first + count + 1, dtype=np.float64)
        return np.stack([k, k, np.full_like(k, 0.5), 0.25 ** k, np.ones_like(k)], axis=1)

    def result(self):
        return np.full((self.job.nz, self.job.m, self.job.n, self.job.n), float(self.enq), np.float32)

    def free(self):
        if not self.freed:
            self.freed = True
            with FakeSolver.lock:
                FakeSolver.live -= 1


def _job(n=16, niter=5, tol=None, fp=10):
    nt, m, nz = 8, 2, 2
    j = Job(np.zeros((nz, nt, n), np.float32), np.zeros(nt, np.float32), np.zeros((m, 2 * nt), np.float32),
            niter=niter, tol=tol)
    j.fp = fp
    return j


def _server(**kw):
    FakeSolver.live = FakeSolver.peak = 0
    fail = kw.pop("fail_names", ())
    return Server(factory=lambda j: FakeSolver(j, fail=j.name in fail), footprint=lambda j: j.fp, **kw)


def test_all_jobs_finish_with_their_iterations():
    srv = _server(budget=100, max_active=4, ahead=2)
    jobs = [srv.submit(_job(niter=k)) for k in (1, 3, 4, 7)]
    srv.run()
    srv.close()
    for j in jobs:
        assert j.error is None and j.iterations == j.niter
        assert float(j.result.flat[0]) == j.niter
        assert j.latency() > 0 and j.t_submit <= j.t_admit <= j.t_start <= j.t_done
    assert FakeSolver.live == 0


def test_admission_respects_memory_budget_and_slots():
    srv = _server(budget=25, max_active=8, ahead=1)
    jobs = [srv.submit(_job(niter=6, fp=10)) for _ in range(6)]
    srv.run()
    srv.close()
    assert all(j.error is None for j in jobs)
    assert FakeSolver.peak == 2          # 2 x 10 bytes fit in 25, 3 do not
    srv = _server(budget=1000, max_active=3, ahead=1)
    jobs = [srv.submit(_job(niter=6, fp=10)) for _ in range(7)]
    srv.run()
    srv.close()
    assert FakeSolver.peak == 3 and srv.peak_active <= 3


def test_continuous_admission_while_others_run():
    srv = _server(budget=1000, max_active=2, ahead=1)
    long_job = srv.submit(_job(niter=40))
    short = [srv.submit(_job(niter=2)) for _ in range(4)]
    srv.run()
    srv.close()
    # the short jobs pass through the second slot while the long one runs
    assert all(j.t_done < long_job.t_done for j in short)


def test_late_arrivals_are_admitted():
    srv = _server(budget=1000, max_active=4, ahead=1)
    first = srv.submit(_job(niter=20))
    late = []

    def producer():
        time.sleep(0.01)
        late.append(srv.submit(_job(niter=2)))

    th = threading.Thread(target=producer)
    th.start()
    srv.run(until=lambda: not th.is_alive() and srv.pending() == 0)
    th.join()
    srv.close()
    assert late and late[0].t_done < first.t_done


def test_early_stop_on_relative_change():
    srv = _server(budget=1000, max_active=2, ahead=1)
    j = srv.submit(_job(niter=50, tol=1e-3))
    srv.run()
    srv.close()
    # relative change 0.5**k < 1e-3 first at k = 10
    assert j.iterations == 10 and j.error is None


def test_failed_job_does_not_stop_others_and_oversized_is_rejected():
    srv = _server(budget=50, max_active=4, ahead=1, fail_names=("bad",))
    ok = srv.submit(_job(niter=3))
    bad = _job(niter=3)
    bad.name = "bad"
    srv.submit(bad)
    big = srv.submit(_job(niter=3, fp=500))
    srv.run()
    srv.close()
    assert ok.error is None and ok.iterations == 3
    assert bad.error and "bad inputs" in bad.error
    assert big.error and "budget" in big.error
    assert srv.used == 0 and FakeSolver.live == 0


def test_autotune_policy():
    """autotune="auto" tunes only jobs with n >= AUTOTUNE_MIN_N; True / False force it."""
    from rectv_gpu_amd import batch

    class Tracking(FakeSolver):
        tuned = []

        def autotune(self):
            Tracking.tuned.append(self.job.n)
            return 8, 4

    small = Job(np.zeros((2, 8, 64), np.float32), np.zeros(8, np.float32), np.zeros((2, 16), np.float32), niter=1)
    big = Job(np.zeros((1, 4, batch.AUTOTUNE_MIN_N), np.float32), np.zeros(4, np.float32),
              np.zeros((2, 8), np.float32), niter=1)
    for job, mode, want in [(small, "auto", []), (big, "auto", [big.n]), (small, True, [small.n]), (big, False, [])]:
        Tracking.tuned = []
        s = batch.prepare(job, factory=lambda j: Tracking(j), graph=False, autotune=mode)
        s.free()
        assert Tracking.tuned == want, (job.n, mode)



def test_enqueue_failure_fails_that_job_only():
    """A solver error while iterating (e.g. a HIP error reported by iterate_async)
    ends that job with its error and frees its memory; the others finish."""

    class Flaky(FakeSolver):
        def iterate_async(self, k, titer, l0, l1):
            if self.job.name == "flaky" and self.enq >= 2:
                raise RuntimeError("device error in enqueue")
            super().iterate_async(k, titer, l0, l1)

    FakeSolver.live = FakeSolver.peak = 0
    srv = Server(factory=lambda j: Flaky(j), footprint=lambda j: j.fp, budget=100, max_active=3, ahead=1)
    ok = [srv.submit(_job(niter=4)) for _ in range(2)]
    fl = _job(niter=6)
    fl.name = "flaky"
    srv.submit(fl)
    srv.run()
    srv.close()
    assert all(j.error is None and j.iterations == 4 for j in ok)
    assert fl.error and "enqueue" in fl.error and fl.done.is_set()
    assert srv.used == 0 and FakeSolver.live == 0
