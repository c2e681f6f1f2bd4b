"""The oracle against the live reference build (oracle/_ref) on degenerate
lattices: extents of one and two cells, open and periodic faces, all fluid,
20 steps with the force on.  aa-soa bit-exact; list-aa-pv-soa adjacency
bit-exact and node views within 1e-14 of the largest value (the reference
takes its vectorised path for runs of consecutive adjacency, whose
association order differs from the scalar tree the oracle follows)."""
import numpy as np
import pytest

from conftest import DEGENERATE_SHAPES


@pytest.mark.parametrize("dims,per", DEGENERATE_SHAPES)
def test_degenerate_shapes_match_reference(oracle, reference, dims, per):
    mask = np.ones(int(np.prod(dims)), dtype=np.int32)
    a, _ = oracle.aa_run(mask, dims, per, 20, 1.3, 1e-5)
    ra, _ = reference.aa_run(mask, dims, per, 20, 1.3, 1e-5)
    assert np.array_equal(a, ra)
    nodes, b = oracle.list_run(mask, dims, per, 20, 1.3, 1e-5)
    rnodes, rb, _ = reference.list_run(mask, dims, per, 20, 1.3, 1e-5)
    assert np.array_equal(b.adj, rb.adj)
    assert np.abs(nodes - rnodes).max() <= 1e-14 * np.abs(rnodes).max()
