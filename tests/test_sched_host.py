"""The server's scheduler (csrc/mcrt_sched.cpp) on the CPU over a simulated
device (tests/native/sched_harness.cpp): admission cap under overload, the
publication of tickets whose launch or window failed (once, after nothing in
flight holds them), preemption and resumption of stopped chunks, cancellation.
The harness checks the scheduler's calls against what the device path relies
on and counts violations; every test requires none."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = [ROOT / "tests" / "native" / "sched_harness.cpp",
       ROOT / "montecarloradiativetransfer.jl_amd" / "csrc" / "mcrt_sched.cpp"]
DONE, CANCELLED, FAILED = 0, 1, 2
MCRT_ERR_NOMEM, MCRT_ERR_HIP = -3, -4
STATS = ["submitted", "completed", "cancelled", "failed", "launches", "tasks", "resumed",
         "preemptions", "finishes", "queued", "running", "in_use", "max_running", "max_in_use",
         "violations"]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("sched") / "libsched_harness.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-Wall", "-o", str(out)] +
                   [str(s) for s in SRC], check=True)
    L = C.CDLL(str(out))
    L.sh_create.restype = C.c_void_p
    L.sh_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
    for n in ("sh_destroy", "sh_cancel_all"):
        getattr(L, n).argtypes = [C.c_void_p]
    L.sh_submit.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_double, C.c_int64, C.c_double,
                            C.c_int, C.c_double]
    L.sh_cancel.argtypes = [C.c_void_p, C.c_uint64]
    L.sh_fail_launches.argtypes = [C.c_void_p, C.c_int]
    L.sh_fail_windows.argtypes = [C.c_void_p, C.c_int]
    L.sh_set_slice.argtypes = [C.c_void_p, C.c_double]
    L.sh_step.argtypes = [C.c_void_p, C.c_double]
    for n in ("sh_idle", "sh_window_live"):
        getattr(L, n).argtypes = [C.c_void_p]
    for n in ("sh_final", "sh_publish_count", "sh_err", "sh_in_window"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_uint64]
    L.sh_time.restype = C.c_double
    L.sh_time.argtypes = [C.c_void_p, C.c_uint64, C.c_int]
    L.sh_stats.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.sh_violation.restype = C.c_char_p
    L.sh_violation.argtypes = [C.c_void_p, C.c_int]
    return L


class H:
    def __init__(self, L, blocks=16, window=32, conts=48, max_active=1024, budget=1e12,
                 slice_ms=2.0, iter_ms=0.05):
        self.L = L
        self.h = L.sh_create(blocks, window, conts, max_active, budget, slice_ms, iter_ms)
        self.now = 0.0

    def submit(self, id, nchunks, chunk_ms, mixed=True, need=1e6, fail_create=False):
        self.L.sh_submit(self.h, id, int(mixed), need, nchunks, chunk_ms, int(fail_create), self.now)

    def step(self, dt=0.01):
        self.now += dt
        return self.L.sh_step(self.h, self.now)

    def run(self, until=None, dt=0.01, limit=10**6):
        for _ in range(limit):
            if until is not None and self.now >= until:
                return
            self.step(dt)
            if until is None and self.L.sh_idle(self.h):
                return
        raise AssertionError("the scheduler did not drain")

    def stats(self):
        a = (C.c_double * 15)()
        self.L.sh_stats(self.h, a)
        return dict(zip(STATS, a))

    def violations(self):
        n = int(self.stats()["violations"])
        return [self.L.sh_violation(self.h, i).decode() for i in range(n)]

    def final(self, id):
        return self.L.sh_final(self.h, id)

    def close(self):
        self.L.sh_destroy(self.h)


def test_overload_admits_at_most_1024(lib):
    h = H(lib, blocks=256, window=512, conts=768, max_active=1024)
    n = 3000
    for i in range(1, n + 1):
        h.submit(i, 2, 0.2)
    h.step()
    assert h.stats()["running"] == 1024  # the cap, not the budget, stops admission
    h.run()
    s = h.stats()
    assert h.violations() == []
    assert s["max_running"] == 1024
    assert s["completed"] == n and s["failed"] == 0
    assert all(lib.sh_publish_count(h.h, i) == 1 and h.final(i) == DONE for i in range(1, n + 1))
    assert s["in_use"] == 0 and s["queued"] == 0 and s["running"] == 0
    h.close()


def test_failed_launch_publishes_each_ticket_once_after_release(lib):
    h = H(lib, blocks=8, window=16, conts=24, slice_ms=1.0)
    # long jobs with chunks stopped in flight when the launches start failing
    for i in range(1, 13):
        h.submit(i, 4, 3.0)
    h.run(until=2.5)
    assert lib.sh_window_live(h.h)
    held = [i for i in range(1, 13) if lib.sh_in_window(h.h, i)]
    assert held
    for i in range(13, 17):
        h.submit(i, 2, 0.3)
    lib.sh_fail_launches(h.h, 1)  # the next window is refused before anything is enqueued
    h.run()
    s = h.stats()
    assert h.violations() == []  # includes: never published while a window or read-back held it
    failed = [i for i in range(1, 17) if h.final(i) == FAILED]
    assert failed and s["failed"] == len(failed)
    assert all(lib.sh_err(h.h, i) == MCRT_ERR_HIP for i in failed)
    assert all(lib.sh_publish_count(h.h, i) == 1 for i in range(1, 17))
    assert all(h.final(i) in (DONE, FAILED) for i in range(1, 17))
    assert s["completed"] + s["failed"] == 16 and s["running"] == 0 and s["in_use"] == 0
    # the server keeps working afterwards
    for i in range(17, 21):
        h.submit(i, 3, 0.5)
    h.run()
    assert all(h.final(i) == DONE for i in range(17, 21))
    assert h.violations() == []
    h.close()


def test_failed_window_fails_its_jobs_once(lib):
    h = H(lib, blocks=8, window=16, conts=24)
    for i in range(1, 9):
        h.submit(i, 3, 2.0)
    h.run(until=0.5)
    in_win = [i for i in range(1, 9) if lib.sh_in_window(h.h, i)]
    lib.sh_fail_windows(h.h, 1)  # the device fails the window in flight
    h.run()
    assert h.violations() == []
    assert all(h.final(i) == FAILED and lib.sh_err(h.h, i) == MCRT_ERR_HIP for i in in_win)
    assert all(lib.sh_publish_count(h.h, i) == 1 for i in range(1, 9))
    h.close()


def test_preemption_lets_short_jobs_through_a_long_one(lib):
    h = H(lib, blocks=16, window=32, conts=48, slice_ms=2.0, iter_ms=0.05)
    h.submit(1, 64, 40.0)  # 64 chunks of 40 ms on 16 workgroups: 160 ms alone
    h.run(until=10.0)
    for i in range(2, 22):
        h.submit(i, 2, 0.5)
        h.run(until=h.now + 3.0)
    h.run()
    s = h.stats()
    assert h.violations() == []
    lat = [lib.sh_time(h.h, i, 3) - lib.sh_time(h.h, i, 0) for i in range(2, 22)]
    # a short job waits at most one slice and one iteration for the device
    assert max(lat) < 2.0 + 0.5 + 1.0, lat
    assert h.final(1) == DONE and s["resumed"] > 0 and s["preemptions"] >= 20
    long_lat = lib.sh_time(h.h, 1, 3) - lib.sh_time(h.h, 1, 0)
    assert 160.0 <= long_lat < 200.0  # the long job still progresses at near full rate
    h.close()


def test_no_preemption_without_waiting_work(lib):
    h = H(lib, blocks=16, window=32, conts=48, slice_ms=0.5)
    h.submit(1, 16, 5.0)
    h.run()
    s = h.stats()
    assert s["preemptions"] == 0 and s["launches"] == 1 and s["resumed"] == 0
    assert h.final(1) == DONE and h.violations() == []
    h.close()


def test_cancel_queued_and_running(lib):
    h = H(lib, blocks=4, window=8, conts=12, max_active=2)
    for i in range(1, 5):
        h.submit(i, 4, 2.0)
    h.run(until=1.0)
    lib.sh_cancel(h.h, 1)  # running, chunks in the window
    lib.sh_cancel(h.h, 4)  # still queued (max_active = 2)
    h.run()
    assert h.violations() == []
    assert h.final(1) == CANCELLED and h.final(4) == CANCELLED
    assert h.final(2) == DONE and h.final(3) == DONE
    assert all(lib.sh_publish_count(h.h, i) == 1 for i in range(1, 5))
    h.close()


def test_create_failure_and_budget(lib):
    h = H(lib, blocks=4, window=8, conts=12, budget=5e6)
    h.submit(1, 2, 1.0, fail_create=True)
    h.submit(2, 2, 1.0, need=6e6)  # more than the whole budget
    h.submit(3, 2, 1.0, need=3e6)
    h.submit(4, 2, 1.0, need=3e6)  # waits for 3 to free its memory
    h.step()
    assert h.stats()["max_in_use"] <= 5e6
    h.run()
    assert h.violations() == []
    assert h.final(1) == FAILED and lib.sh_err(h.h, 1) == MCRT_ERR_NOMEM
    assert h.final(2) == FAILED and lib.sh_err(h.h, 2) == MCRT_ERR_NOMEM
    assert h.final(3) == DONE and h.final(4) == DONE
    assert lib.sh_time(h.h, 4, 1) >= lib.sh_time(h.h, 3, 3)  # admitted after 3 finished
    h.close()


def test_precisions_alternate(lib):
    h = H(lib, blocks=4, window=8, conts=12, slice_ms=1.0)
    h.submit(1, 8, 5.0, mixed=True)
    h.submit(2, 8, 5.0, mixed=False)
    h.run()
    assert h.final(1) == DONE and h.final(2) == DONE and h.violations() == []
    # both finish within about the time of running both back to back
    assert max(lib.sh_time(h.h, i, 3) for i in (1, 2)) < 2 * 10.0 + 5.0
    h.close()


def test_randomised_run_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "sched_main"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                        "-fno-omit-frame-pointer", "-DSCHED_HARNESS_MAIN", "-o", str(exe)] +
                       [str(s) for s in SRC], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower():
        pytest.skip("sanitizer runtime not available")
    assert r.returncode == 0, r.stderr
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                       env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "halt_on_error=1"})
    assert p.returncode == 0, p.stdout + p.stderr
    assert "violations 0" in p.stdout and "unpublished 0" in p.stdout
