"""Serving host logic on the CPU: the seeded job stream, first-fit admission
by slots and memory footprint, and the footprint bound."""
import collections
import math

import pytest

from pfm import serving
from pfm.serving import Admission, Job


def test_job_stream_is_seeded_and_in_range():
    a = serving.job_stream(400, rate=50.0, seed=3)
    b = serving.job_stream(400, rate=50.0, seed=3)
    assert a == b
    assert a != serving.job_stream(400, rate=50.0, seed=4)
    arr = [j.arrival for j in a]
    assert all(x < y for x, y in zip(arr, arr[1:]))
    # Poisson arrivals: mean gap 1/rate
    assert abs(arr[-1] / len(arr) - 1 / 50.0) < 0.3 / 50.0
    n2 = sum(j.dim == 2 for j in a)
    assert 0.6 < n2 / len(a) < 0.8
    for j in a:
        n = j.res[0]
        assert all(r == n for r in j.res)
        assert n % 8 == 0
        lo, hi = (128, 1024) if j.dim == 2 else (64, 192)
        assert lo <= n <= hi
        assert 20 <= j.steps <= 200
        assert j.ppc == (4 if j.dim == 2 else 8)
    # log-uniform: the median size is near the geometric mean of the range
    med2 = sorted(j.res[0] for j in a if j.dim == 2)[n2 // 2]
    assert 250 < med2 < 480


def _jobs(sizes):
    return [Job(id=i, res=(8, 8), steps=1, seed=i, arrival=0.0) for i in range(len(sizes))], dict(enumerate(sizes))


def test_admission_respects_slots():
    jobs, fp = _jobs([1] * 5)
    adm = Admission(100, slots=2, footprint=lambda j: fp[j.id])
    pending = collections.deque(jobs)
    got, rej = adm.admit_from(pending)
    assert [j.id for j in got] == [0, 1] and not rej
    assert [j.id for j in pending] == [2, 3, 4]
    assert adm.admit_from(pending) == ([], [])
    adm.release(got[0])
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [2]


def test_admission_first_fit_by_memory():
    # a large job at the head waits; smaller ones behind it that fit go first
    jobs, fp = _jobs([60, 50, 30, 10])
    adm = Admission(100, slots=8, footprint=lambda j: fp[j.id])
    pending = collections.deque(jobs)
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [0, 2, 3] and adm.used == 100
    assert [j.id for j in pending] == [1]
    adm.release(got[0])
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [1] and adm.used == 90


def test_admission_rejects_what_never_fits():
    jobs, fp = _jobs([150, 10])
    adm = Admission(100, slots=4, footprint=lambda j: fp[j.id])
    pending = collections.deque(jobs)
    got, rej = adm.admit_from(pending)
    assert [j.id for j in rej] == [0] and [j.id for j in got] == [1] and not pending


def test_admission_retained_bytes_count_until_freed():
    jobs, fp = _jobs([60, 60])
    adm = Admission(100, slots=4, footprint=lambda j: fp[j.id])
    pending = collections.deque(jobs)
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [0]
    adm.release(got[0], retain=50)
    assert adm.used == 0 and adm.retained == 50
    assert adm.admit_from(pending) == ([], [])
    assert adm.blocked_by_retained(pending)
    adm.free_retained()
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [1]
    with pytest.raises(ValueError):
        Admission(1, slots=0)


def test_admission_batch_memory_counted_until_the_batch_is_dropped():
    """A batch of two 2D jobs (no slot) finishes while a slot job runs.  Its
    jobs are released together with their whole footprints retained, so a job
    that needs that memory waits until the server frees the batch."""
    jobs, fp = _jobs([30, 30, 30, 60])
    batched = {0, 1}
    adm = Admission(100, slots=1, footprint=lambda j: fp[j.id], takes_slot=lambda j: j.id not in batched)
    pending = collections.deque(jobs)
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [0, 1, 2] and adm.used == 90 and adm.slots_used == 1
    adm.release_batch(got[:2])
    assert adm.used == 30 and adm.retained == 60
    # job 3 (60) fits in the budget only once the batch's 60 retained bytes are freed
    assert adm.admit_from(pending) == ([], [])
    assert not adm.blocked_by_retained(pending)  # the slot is still taken by job 2
    adm.release(got[2], retain=0)
    assert adm.blocked_by_retained(pending)
    assert adm.admit_from(pending) == ([], [])
    adm.free_retained()
    got, _ = adm.admit_from(pending)
    assert [j.id for j in got] == [3]


def test_footprint_bounds_the_allocations():
    for res, prec in (((64, 64), 1), ((128, 128), 0), ((32, 32, 32), 1), ((48, 48, 48), 0)):
        job = Job(id=0, res=res, steps=1, seed=0)
        d, cells = len(res), math.prod(res)
        n = cells * job.ppc
        fp = serving.footprint_bytes(job, prec)
        # the particle state alone: 3D 132 B/particle (imp_grad_m not kept), 2D 112 B
        assert fp > n * (132 if d == 3 else 112)
        assert fp > serving.mg_bytes(res, 3, prec) + serving.binning_bytes(res, n)
    small = serving.footprint_bytes(Job(0, (64, 64, 64), 1, 0), 1)
    big = serving.footprint_bytes(Job(0, (128, 128, 128), 1, 0), 1)
    assert 6 < big / small < 9
    assert serving.mg_bytes((128,) * 3, 3, 1) < serving.mg_bytes((128,) * 3, 3, 0)
