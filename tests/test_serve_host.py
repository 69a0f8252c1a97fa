"""Host logic of the served stream (no GPU): admission only of jobs that have
arrived, only into free slots, only while the resident and prepared boxes'
footprint fits the memory budget; slots are refilled as jobs finish.  The device batch is replaced
by a stand-in that advances each slot by a fixed number of cycles per launch."""
import numpy as np

from graspa_amd import serve, synth


class _FakeBox:
    def __init__(self, system, precision=1, device=0, capacity=1024):
        self.system = system
        self.capacity = capacity

    def init_ewald(self):
        return None

    def footprint(self):
        return serve.footprint_bytes(self.system, self.capacity)

    def close(self):
        pass


class _FakeBatch:
    log = []

    def __init__(self, boxes, seeds, streams=None, trace_cap=0, cycle_cap=0, n_slots=None):
        self.n = n_slots
        self.cyc = np.zeros(n_slots, dtype=np.int64)
        self.tgt = np.zeros(n_slots, dtype=np.int64)
        self.on = [False] * n_slots

    def set_box(self, slot, box, params=None, seed=0, stream=0):
        self.on[slot] = box is not None
        self.cyc[slot] = 0
        self.tgt[slot] = 0

    def set_cycles(self, slot, n, production=True):
        self.tgt[slot] = self.cyc[slot] + n

    def launch_async(self, mpl):
        _FakeBatch.log.append(sum(self.on))
        for i in range(self.n):
            if self.on[i] and self.cyc[i] < self.tgt[i]:
                self.cyc[i] += 1

    def wait(self):
        return int(sum(1 for i in range(self.n) if self.on[i] and self.cyc[i] < self.tgt[i]))

    def total_moves(self):
        return int(self.cyc.sum())

    def slot_done(self, i):
        return self.on[i] and self.cyc[i] >= self.tgt[i]

    def stats(self, i):
        return int(self.cyc[i])

    def cycle_times(self, i, n):
        return np.ones(int(self.cyc[i]))

    def close(self):
        pass


def test_admission_respects_slots_budget_and_arrivals(monkeypatch):
    monkeypatch.setattr(serve.engine, "Box", _FakeBox)
    monkeypatch.setattr(serve.engine, "Batch", _FakeBatch)
    monkeypatch.setattr(serve.engine, "make_mc_params", lambda s, b: None)
    systems = [synth.build(n_fw=100 + 50 * j, L=24.0, species=("CO2",), seed=j) for j in range(6)]
    fp = [serve.footprint_bytes(s, 64) for s in systems]
    budget = fp[0] + fp[1] + max(fp) // 2       # never three of the larger jobs at once
    jobs = [serve.Job(job_id=j, system=systems[j], cycles=3 + j, seed=j, arrival=0.0, capacity=64)
            for j in range(6)]
    _FakeBatch.log = []
    srv = serve.Server(n_slots=3, memory_budget=budget, moves_per_launch=10)
    done = srv.run(jobs)
    assert sorted(j.job_id for j in done) == list(range(6))
    for j in done:
        assert j.stats == j.cycles
        assert j.admitted <= j.finished
    # at most two jobs resident at once under this budget, and some launches ran two
    assert max(_FakeBatch.log) == 2 and srv.max_concurrent == 2
    assert srv.resident == 0 and srv.prepared == 0 and not srv.ready
    # the resident plus the prepared boxes never exceeded the budget
    assert srv.peak_bytes <= budget
    # jobs were admitted in arrival order
    order = sorted(done, key=lambda j: j.admitted)
    assert [j.job_id for j in order] == sorted(j.job_id for j in order)


def test_footprint_grows_with_framework_and_k_vectors():
    a = synth.build(n_fw=500, L=24.0)
    b = synth.build(n_fw=5000, L=24.0)
    c = synth.build(n_fw=500, L=40.0)
    assert serve.footprint_bytes(b) > serve.footprint_bytes(a)
    assert serve.footprint_bytes(c) > serve.footprint_bytes(a)
