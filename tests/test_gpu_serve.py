"""Served stream with continuous batching: jobs admitted into batch slots as
slots free up, sharing the device with jobs already running; every job's
result must be the same as when it runs alone (bit-identical here, since a
job's moves depend only on its own counter-based stream)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _job_system(j):
    from graspa_amd import synth
    return synth.build(n_fw=150 + 40 * j, L=24.0 + 0.5 * j, species=("CO2", "N2"), mol_fraction=[0.5, 0.5],
                       seed=500 + j, T=280.0 + 10 * j, P=10 ** (5.5 + 0.25 * j), cycles=10,
                       probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.3))


def _alone(engine_mod, job, prec):
    box = engine_mod.Box(job.system, precision=prec, capacity=job.capacity)
    box.init_ewald()
    bt = engine_mod.Batch([box], [job.seed], [job.stream], trace_cap=0, cycle_cap=64)
    bt.run(job.cycles, production=True)
    st = bt.stats(0)
    bt.close()
    box.close()
    return st


@pytest.mark.parametrize("prec,slice_us", [(1, 0.0), (1, 200.0), (0, 200.0)])
def test_served_jobs_equal_alone(engine_mod, prec, slice_us):
    from graspa_amd import serve
    jobs = [serve.Job(job_id=j, system=_job_system(j), cycles=4 + 3 * j, seed=900 + j, stream=j,
                      arrival=0.002 * j, capacity=128) for j in range(5)]
    print("served case jobs=5 slots=2 precision=%d slice=%g" % (prec, slice_us), flush=True)
    srv = serve.Server(n_slots=2, precision=prec, moves_per_launch=23, trace_cap=0, cycle_cap=64,
                       time_slice_us=slice_us)
    done = srv.run(jobs)
    srv.close()
    assert sorted(j.job_id for j in done) == list(range(5))
    # launches ran two jobs at once (continuous batching, not a queue)
    assert srv.max_concurrent == 2
    if slice_us == 0.0:
        assert srv.launches < sum(j.stats.moves for j in done) / 23
    for j in done:
        ref = _alone(engine_mod, serve.Job(job_id=j.job_id, system=_job_system(j.job_id), cycles=j.cycles,
                                           seed=j.seed, stream=j.stream, capacity=128), prec)
        st = j.stats
        assert st.cycles == j.cycles == ref.cycles
        assert st.moves == ref.moves
        assert np.array_equal(st.n_mol, ref.n_mol)
        assert np.array_equal(st.counters, ref.counters)
        for k in ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald"):
            assert getattr(st.running, k) == getattr(ref.running, k), (j.job_id, k)
        assert np.array_equal(st.widom_w, ref.widom_w)
        assert len(j.cycle_seconds) == j.cycles and np.all(j.cycle_seconds > 0)


def test_box_footprint_reports_device_memory(engine_mod):
    """graspa_box_footprint counts the box's device allocations: at least the
    framework copies and the guest store, growing when a Widom batch grows the
    pipeline's buffers, and the same for a box built again from the same job."""
    from graspa_amd import synth
    s = synth.config1(n_fw=750)
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_MIXED, capacity=64)
    box.init_ewald()
    f0 = box.footprint()
    n_fw = len(s.fw_type)
    assert f0 >= n_fw * (32 + 4 + 16 + 16 + 32 + 4) + 64 * 3 * (32 + 12 + 16)
    box.widom_batch(1, 20000, seed=3, results=False)
    f1 = box.footprint()
    assert f1 > f0 + 20000 * 10 * 16
    box.close()
    again = engine_mod.Box(s, precision=engine_mod.GRASPA_MIXED, capacity=64)
    again.init_ewald()
    assert again.footprint() == f0
    again.close()
