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


@pytest.mark.gpu
def test_register_while_requests_are_in_flight():
    # a cluster and a periodic system registered while the first clusters' requests run; every result matches the
    # oracle, and the earlier systems' results match their results before the registration
    catalog = cluster_catalog([20, 45, 6], seed=11)
    server = ih.Server("mixed", max_batch=8)
    entries = [server.add_system(s) for s, _ in catalog]
    server.start()
    before = [server.wait(server.submit(e, p), e) for e, (_, p) in zip(entries, catalog)]
    tickets = [(server.submit(entries[k % 3], catalog[k % 3][1]), k % 3) for k in range(60)]
    new_system, new_pos = workloads.water_cluster(33, seed=99)
    added = server.add_system(new_system)
    box_system, box_pos = workloads.alchemical_box(0.3, num_waters=300, box=2.1, cutoff=0.9)
    alch = server.add_system(box_system, box_pos)
    t_new = server.submit(added, new_pos)
    t_alch = server.submit(alch, box_pos, 0.6)
    more = [(server.submit(entries[k % 3], catalog[k % 3][1]), k % 3) for k in range(20)]
    for t, k in tickets+more:
        energy, forces, _ = server.wait(t, entries[k])
        assert energy == before[k][0] and np.array_equal(forces, before[k][1])
    energy, forces, _ = server.wait(t_new, added)
    e_ref, f_ref, scale = oracle.evaluate(new_system, new_pos, with_scale=True)
    assert abs(energy-e_ref) <= 2e-6*max(scale, abs(e_ref))
    assert rms_rel(forces, f_ref) <= 1e-5
    energy, forces, _ = server.wait(t_alch, alch)
    box = box_system.getDefaultPeriodicBoxVectors()
    e_ref, f_ref, scale = oracle.evaluate(box_system, box_pos, box, parameters={"lambda_sterics": 0.6}, with_scale=True)
    assert abs(energy-e_ref) <= 2e-6*max(scale, abs(e_ref))
    assert rms_rel(forces, f_ref) <= 1e-5
    server.close()


@pytest.mark.gpu
def test_several_requests_of_one_system_in_a_batch():
    # rounds of one request per system: the same system with different coordinates in one batch, each result equal to
    # the same request alone
    catalog = cluster_catalog([30, 8], seed=21)
    server = ih.Server("mixed", max_batch=32)
    entries = [server.add_system(s) for s, _ in catalog]
    server.start()
    rng = np.random.default_rng(3)
    coords = [catalog[0][1]+rng.normal(0, 0.01, size=catalog[0][1].shape) for _ in range(6)]
    alone = [server.wait(server.submit(entries[0], c), entries[0]) for c in coords]
    b0, r0 = server.statistics()
    tickets = [server.submit(entries[0], c) for c in coords]+[server.submit(entries[1], catalog[1][1])]
    results = [server.wait(t, entries[0]) for t in tickets[:-1]]
    server.wait(tickets[-1], entries[1])
    b1, r1 = server.statistics()
    assert r1-r0 == 7 and b1-b0 < 7
    for (e1, f1, _), (e2, f2, _) in zip(results, alone):
        assert e1 == e2 and np.array_equal(f1, f2)
    server.close()


@pytest.mark.gpu
def test_window_requests():
    # the energies of one configuration at 16 lambda values through the fused evaluation, against single-lambda
    # requests and the oracle
    system, pos = workloads.alchemical_box(0.3, num_waters=300, box=2.1, cutoff=0.9, seed=4, use_lrc=True)
    server = ih.Server("mixed", max_batch=8)
    alch = server.add_system(system, pos)
    server.start()
    values = np.linspace(0, 1, 16)
    energies, ms = server.wait_windows(server.submit_windows(alch, pos, values))
    single = np.array([server.wait(server.submit(alch, pos, float(v)), alch)[0] for v in values])
    assert ms > 0
    assert np.abs(energies-single).max() <= 2e-6*np.abs(single).max()
    assert abs(energies[-1]-energies[0]) > 1.0
    box = system.getDefaultPeriodicBoxVectors()
    for v, e in zip(values[::5], energies[::5]):
        ref = oracle.evaluate(system, pos, box, parameters={"lambda_sterics": float(v)})[0]
        assert abs(e-ref) <= 2e-5*max(abs(ref), 1.0), (v, e, ref)
    server.close()


@pytest.mark.gpu
def test_memory_is_measured_and_bounded():
    # cluster requests: their staging bytes in flight never exceed the budget
    catalog = cluster_catalog([12, 70, 25, 40], seed=31)
    per_request = 56*max(s.getNumParticles() for s, _ in catalog)+24
    server = ih.Server("mixed", max_batch=64, memory_budget=2*per_request)
    entries = [server.add_system(s) for s, _ in catalog]
    server.start()
    assert server.memory()["resident_bytes"] > 0   # the two lanes' catalogs
    tickets = [(server.submit(e, p), e) for e, (_, p) in zip(entries, catalog) for _ in range(4)]
    for t, e in tickets:
        server.wait(t, e)
    memory = server.memory()
    assert memory["in_flight_bytes"] == 0
    assert 0 < memory["peak_in_flight_bytes"] <= 2*per_request
    server.close()
    # periodic requests hold their slot's device memory, measured at registration: with a budget of one byte two
    # requests are never in flight together
    box_system, box_pos = workloads.alchemical_box(0.0, num_waters=300, box=2.1, cutoff=0.9)
    server = ih.Server("mixed", max_batch=64, memory_budget=1)
    alch = server.add_system(box_system, box_pos)
    server.start()
    slot_bytes = server.memory()["resident_bytes"]//4   # four slots per periodic system
    assert slot_bytes > 56*box_system.getNumParticles()
    tickets = [server.submit(alch, box_pos, 0.5) for _ in range(4)]
    for t in tickets:
        server.wait(t, alch)
    assert server.memory()["peak_in_flight_bytes"] == slot_bytes
    server.close()
