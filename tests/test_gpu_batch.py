"""Batched evaluation of independent clusters (the batched configuration) against the oracle, and the properties the
batch promises: every System's result is the same alone and in a batch, and graph replays equal eager launches."""
import numpy as np
import pytest

import interax_hip as ih
import oracle
import workloads

# mixed precision: the reference's single/mixed export tolerances (TestInteraxExportValues.h:195-196, energy 2e-6
# relative, forces 5e-5 of the RMS force) and the 1e-5 force RMS target; double precision against the FP64 oracle
ENERGY_TOL = {"mixed": 2e-6, "double": 1e-9}
FORCE_TOL = {"mixed": 5e-5, "double": 1e-8}


def clusters(sizes, seed=0):
    out = [workloads.water_cluster(m, seed=seed+k) for k, m in enumerate(sizes)]
    return [s for s, _ in out], [p for _, p in out]


def check_against_oracle(systems, positions, energies, forces, precision):
    for system, pos, e, f in zip(systems, positions, energies, forces):
        e_ref, f_ref, scale = oracle.evaluate(system, pos, with_scale=True)
        assert abs(e-e_ref) <= ENERGY_TOL[precision]*max(scale, abs(e_ref)), (e, e_ref, scale)
        rms = np.sqrt(np.mean(f_ref**2))
        assert np.abs(f-f_ref).max() <= FORCE_TOL[precision]*rms+1e-8
        if precision == "mixed":
            assert np.linalg.norm(f-f_ref) <= 1e-5*np.linalg.norm(f_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_batch_against_oracle(precision):
    systems, positions = clusters([2, 5, 13, 1, 40, 22, 64, 3])
    batch = ih.Batch(systems, precision)
    energies, forces = batch.evaluate(positions)
    check_against_oracle(systems, positions, energies, forces, precision)
    batch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_large_clusters_against_oracle(precision):
    # sites with more than 64 later neighbours take the triple kernel's blocks of 64 neighbours (full blocks with the
    # forces summed in registers, a partial last block), and the 200-molecule cluster's first sites (more than
    # 16 384 neighbour pairs) are split into several work items
    systems, positions = clusters([200, 130, 66], seed=11)
    batch = ih.Batch(systems, precision)
    energies, forces = batch.evaluate(positions)
    check_against_oracle(systems, positions, energies, forces, precision)
    batch.close()


@pytest.mark.gpu
def test_same_result_alone_and_in_a_batch():
    systems, positions = clusters([7, 30, 2, 90, 11], seed=5)
    batch = ih.Batch(systems, "mixed")
    energies, forces = batch.evaluate(positions)
    for k in (0, 3, 4):
        alone = ih.Batch([systems[k]], "mixed")
        e1, f1 = alone.evaluate([positions[k]])
        assert e1[0] == energies[k]
        assert np.array_equal(f1[0], forces[k])
        alone.close()
    batch.close()


@pytest.mark.gpu
def test_graph_replays_equal_eager_launches():
    systems, positions = clusters([12, 3, 48, 20], seed=9)
    graph = ih.Batch(systems, "mixed")
    eager = ih.Batch(systems, "mixed")
    eager.set_graph(False)
    rng = np.random.default_rng(2)
    for step in range(5):
        pos = [p+rng.normal(0, 0.003, size=p.shape) for p in positions]
        e_g, f_g = graph.evaluate(pos)
        e_e, f_e = eager.evaluate(pos)
        assert np.array_equal(e_g, e_e), step
        for a, b in zip(f_g, f_e):
            assert np.array_equal(a, b), step
    graph.close()
    eager.close()


@pytest.mark.gpu
def test_autotune_keeps_results():
    systems, positions = clusters([16, 40, 5, 64], seed=3)
    batch = ih.Batch(systems, "mixed")
    e0, f0 = batch.evaluate(positions)
    chosen = batch.autotune()
    assert chosen["pair_workgroups"] > 0 and chosen["triple_waves"] in (1, 2, 4, 8)
    assert chosen == batch.shape()
    e1, f1 = batch.evaluate(positions)
    assert np.array_equal(e0, e1)
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)
    batch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_every_launch_shape_gives_the_same_results(precision):
    # the shares of a site's neighbour pairs depend on its list capacity only, and are summed in fixed point (forces)
    # or in share order (energies): the wavefronts per work item and the pairs per work item change nothing
    systems, positions = clusters([70, 3, 130, 9, 41], seed=5)
    batch = ih.Batch(systems, precision)
    reference = None
    try:
        for workgroups, waves, pairs in [(4096, 4, 16384), (256, 1, 1024), (512, 2, 8192), (1024, 8, 65536), (2048, 4, 4096),
                                         (8192, 8, 1024), (4096, 1, 262144)]:
            print(f"shape: {workgroups} pair workgroups, {waves} waves, {pairs} pairs per item", flush=True)
            batch.set_shape(workgroups, waves, pairs)
            e, f = batch.evaluate(positions)
            e_graph, f_graph = batch.evaluate(positions)
            assert np.array_equal(e, e_graph)
            if reference is None:
                reference = (e, f)
                continue
            assert np.array_equal(e, reference[0])
            for a, b in zip(f, reference[1]):
                assert np.array_equal(a, b)
    finally:
        batch.close()


def damped_cluster(num_molecules, seed):
    """A water-like cluster whose ATM force has Tang-Toennies and centroid damping, a second (two-particle) site on
    every fourth molecule and a cutoff inside the cluster, so that the segmented triple kernel's damped form, its
    molecule filter and its cutoff test all take part."""
    rng = np.random.default_rng(seed)
    radius = (3*num_molecules/(4*np.pi*33.4))**(1.0/3.0)
    centers = workloads.rsa_points(rng, num_molecules, 2*radius+2.0, 0.25, periodic=False, sphere_radius=radius)
    positions = workloads.water_sites(rng, centers)
    system = ih.System(len(positions))
    workloads.add_water_forces(system, num_molecules, rng, ih.InteraxPairForce.NoCutoff, 1.0)
    tb = ih.InteraxThreeBodyForce()
    for m in range(num_molecules):
        tb.addSite([3*m], 1.45e-3, 1150.0, 0.05+0.01*(m % 3), m)
        if m % 4 == 0:
            tb.addSite([3*m+1, 3*m+2], 0.4e-3, 1300.0, 0.04, m)
    tb.setDampingMethod(ih.InteraxThreeBodyForce.TangToennies)
    tb.setDampingScale(0.9)
    tb.setCentroidDampingRadius(0.2)
    tb.setCutoffDistance(1.0)
    system.addForce(tb)
    return system, positions


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_damped_sites_against_oracle(precision):
    # more than 64 sites per cluster (full and partial neighbour blocks) with damping, sites of one molecule and
    # pairs beyond the cutoff; the same results for every launch shape
    out = [damped_cluster(n, seed=20+k) for k, n in enumerate([150, 9, 70])]
    systems, positions = [s for s, _ in out], [p for _, p in out]
    batch = ih.Batch(systems, precision)
    try:
        energies, forces = batch.evaluate(positions)
        check_against_oracle(systems, positions, energies, forces, precision)
        for workgroups, waves, pairs in [(256, 1, 1024), (1024, 8, 65536), (512, 2, 8192)]:
            print(f"shape: {workgroups} pair workgroups, {waves} waves, {pairs} pairs per item", flush=True)
            batch.set_shape(workgroups, waves, pairs)
            e, f = batch.evaluate(positions)
            assert np.array_equal(e, energies)
            for a, b in zip(f, forces):
                assert np.array_equal(a, b)
    finally:
        batch.close()


@pytest.mark.gpu
def test_batch_rejects_cutoff_methods():
    system, pos = workloads.config0()
    with pytest.raises(ih.InteraxError, match="NoCutoff"):
        ih.Batch([system])
