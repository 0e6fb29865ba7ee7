"""The numpy workload generators used by bench.py match the pinned oracle."""
import numpy as np
import pytest

from amgt_amd import workloads as W
from oracle import oracle as O


@pytest.mark.parametrize("kind,dims,c", [(7, (9, 7, 5), (1, 1, 1)), (7, (12, 9, 1), (1, 1, 1)),
                                         (7, (6, 5, 7), (1.0, 1.0, 0.01)), (9, (8, 5, 1), (1, 1, 1)),
                                         (27, (5, 6, 4), (1, 1, 1))])
def test_stencil_matches_oracle(kind, dims, c):
    a = W.stencil(kind, *dims, c=c)
    b = O.stencil(kind, *dims, c=c)
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and np.array_equal(x, y)


def test_rng_matches_oracle():
    assert np.array_equal(W.pm_stream(1, 10000), O.pm_sequence(1, 10000))
    assert np.array_equal(W.pm_stream(0, 100), O.pm_sequence(0, 100))
    assert np.array_equal(W.rand_uniform(0, 3001), O.rand_uniform(0, 3001))
    # numpy's and libm's log/cos/sin may differ in the last ulp
    assert np.allclose(W.rand_normal(2, 3001), O.rand_normal(2, 3001), rtol=1e-14, atol=1e-15)
