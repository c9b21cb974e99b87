"""The serving loop: results against the oracle, the same result alone and in a mix of concurrent requests,
admission by the memory budget, and the scheduled replay used for latency measurements."""
import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads


def cluster_catalog(sizes, seed=0):
    return [workloads.water_cluster(m, seed=seed+k) for k, m in enumerate(sizes)]


def rms_rel(f, f_ref):
    return np.linalg.norm(f-f_ref)/np.linalg.norm(f_ref)


@pytest.mark.gpu
def test_served_results_against_oracle():
    catalog = cluster_catalog([3, 17, 40, 8])
    box_system, box_pos = workloads.alchemical_box(0.3, num_waters=300, box=2.1, cutoff=0.9)
    server = ih.Server("mixed", max_batch=8)
    entries = [server.add_system(s) for s, _ in catalog]
    alch = server.add_system(box_system)
    server.start()
    tickets = [(server.submit(e, p), e, s, p) for e, (s, p) in zip(entries, catalog)]
    lam = 0.4
    t_alch = server.submit(alch, box_pos, lam)
    for t, e, s, p in tickets:
        energy, forces, ms = server.wait(t, e)
        e_ref, f_ref, scale = oracle.evaluate(s, p, with_scale=True)
        assert abs(energy-e_ref) <= 2e-6*max(scale, abs(e_ref))
        assert rms_rel(forces, f_ref) <= 1e-5
        assert ms > 0
    energy, forces, ms = server.wait(t_alch, alch)
    box = box_system.getDefaultPeriodicBoxVectors()
    e_ref, f_ref, scale = oracle.evaluate(box_system, box_pos, box, parameters={"lambda_sterics": lam}, with_scale=True)
    assert abs(energy-e_ref) <= 2e-6*max(scale, abs(e_ref))
    assert rms_rel(forces, f_ref) <= 1e-5
    server.close()


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix():
    catalog = cluster_catalog([5, 60, 2, 25, 90, 12], seed=4)
    server = ih.Server("mixed", max_batch=16)
    entries = [server.add_system(s) for s, _ in catalog]
    server.start()
    target = 1
    alone_e, alone_f, _ = server.wait(server.submit(entries[target], catalog[target][1]), entries[target])
    rng = np.random.default_rng(0)
    tickets = []
    for k in range(40):
        e = int(rng.integers(len(entries)))
        tickets.append((server.submit(entries[e], catalog[e][1]), e))
        if k == 20:
            tickets.append((server.submit(entries[target], catalog[target][1]), target))
    batches, requests = server.statistics()
    for t, e in tickets:
        energy, forces, _ = server.wait(t, entries[e])
        if e == target:
            assert energy == alone_e
            assert np.array_equal(forces, alone_f)
    batches, requests = server.statistics()
    assert requests == 42 and batches < requests   # requests were batched together
    server.close()


@pytest.mark.gpu
def test_memory_budget_limits_batches():
    catalog = cluster_catalog([10, 11, 12, 13], seed=8)
    server = ih.Server("mixed", max_batch=64, memory_budget=1)   # every batch holds one request
    entries = [server.add_system(s) for s, _ in catalog]
    server.start()
    tickets = [(server.submit(e, p), e) for e, (_, p) in zip(entries, catalog) for _ in range(3)]
    for t, e in tickets:
        server.wait(t, e)
    batches, requests = server.statistics()
    assert batches == requests == 12
    server.close()


@pytest.mark.gpu
def test_replay_matches_direct_results():
    catalog = cluster_catalog([4, 30, 9], seed=2)
    server = ih.Server("mixed", max_batch=32)
    entries = [server.add_system(s) for s, _ in catalog]
    server.start()
    direct = [server.wait(server.submit(e, p), e)[0] for e, (_, p) in zip(entries, catalog)]
    rng = np.random.default_rng(5)
    n = 200
    arrival = np.cumsum(rng.exponential(1/5000.0, size=n))
    which = rng.integers(0, len(entries), size=n)
    latency, energies, makespan = server.replay(arrival, which, [catalog[k][1] for k in which])
    assert np.all(latency > 0) and makespan >= arrival[-1]
    for k in range(n):
        assert energies[k] == direct[which[k]]
    server.close()
