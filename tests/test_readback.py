"""Read-back of results larger than the library's pinned staging buffer (GPU).

Result arrays above 16 MB come back through two 128 MB pinned buffers, the DMA of one
chunk overlapping the host copy of the previous one.  A 4e7-photon run whose detection
records (about 190 MB) and pathlength rows (about 145 MB) each span two chunks must return
exactly the oracle's records, in order, with rows that match them.
"""
import numpy as np
import pytest

import fixtures as F
import oracle_ffi as O
from mcml_gpu_amd import make_opts

SEED = 5


@pytest.mark.gpu
def test_large_readback_matches_oracle(ctx):
    sim = F.hemisphere(40_000_000)
    print("readback: hemisphere 4e7 photons", flush=True)
    res = ctx.run(sim, make_opts(replay=True, pathlength=True), seed=SEED)
    stage = 128 << 20
    assert res.records.nbytes > stage and res.pathlength.nbytes > stage
    rec, _ = O.first_pass(O.derive(sim), SEED, 0, sim.number_photons)
    got = res.records[np.lexsort((res.records["sds"], res.records["key"]))]
    assert np.array_equal(got["key"], rec["key"])
    assert np.array_equal(got["weight"].view(np.uint32), rec["weight"].view(np.uint32))
    # rows are in the records' (file) order: the weight column is the records' weights,
    # the step count is at least one, and the pathlength is positive
    assert np.array_equal(res.pathlength[:, 0].view(np.uint32), res.records["weight"].view(np.uint32))
    assert (res.pathlength[:, -1] >= 1).all() and (res.pathlength[:, 1] > 0).all()
