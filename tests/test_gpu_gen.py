"""Device R-MAT generator (sg_gen_rmat_device): the same graph as the host
generator sg_gen_rmat, vertex for vertex and edge for edge."""
import numpy as np
import pytest

import sgraph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,ef,permute,abc", [(12, 8, False, (0.57, 0.19, 0.19)), (12, 8, True, (0.57, 0.19, 0.19)),
                                                  (16, 16, False, (0.57, 0.19, 0.19)), (14, 4, True, (0.45, 0.15, 0.15)),
                                                  (1, 3, False, (0.57, 0.19, 0.19))])
def test_device_graph_equals_host_graph(scale, ef, permute, abc):
    a, b, c = abc
    print(f"[gpu] gen scale={scale} ef={ef} permute={permute} abc={abc}", flush=True)
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=33, a=a, b=b, c=c, permute=permute)
    nv2, row2, col2 = sgraph.gen_rmat_device(scale, ef, seed=33, a=a, b=b, c=c, permute=permute)
    assert nv2 == nv and np.array_equal(row2, row) and np.array_equal(col2, col)
