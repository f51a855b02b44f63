"""b3d_deduplicate_noalloc: the no-allocation form of the GPU deduplication
gives b3d_deduplicate's outputs bit for bit (integer work), on two calls in a
row that share the device scratch, the second with a smaller input."""
import numpy as np
import pytest

from aspartik_amd import B3hError
from aspartik_amd.data import deduplicate

pytestmark = pytest.mark.gpu


def _msa(S, N, pool, seed):
    """N sites drawn from `pool` distinct random columns (so most repeat),
    DnaNucleotide bytes including gaps (0) and N (15)."""
    rng = np.random.default_rng(seed)
    cols = rng.choice(np.array([0, 1, 2, 4, 8, 15, 5], np.uint8), size=(pool, S),
                      p=[0.02, 0.24, 0.24, 0.24, 0.22, 0.02, 0.02])
    return np.ascontiguousarray(cols[rng.integers(0, pool, size=N)].T)


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_noalloc_equals_allocating_form_on_consecutive_calls(hip_device):
    big = _msa(150, 60_000, 5_000, seed=1)
    small = _msa(40, 5_000, 700, seed=2)
    print("[no-alloc] 150 x 60000, then 40 x 5000", flush=True)
    got_big = deduplicate(big, device=0, no_alloc=True)
    got_small = deduplicate(small, device=0, no_alloc=True)
    _same(got_big, deduplicate(big, device=0))
    _same(got_small, deduplicate(small, device=0))
    # and the host path (the first-occurrence order, counts, site map)
    _same(got_small, deduplicate(small))
    assert got_big[0].shape[1] < 5_001 and got_small[0].shape[1] < 701
    # the scratch keeps the previous call's contents: a repeat is unchanged
    _same(deduplicate(small, device=0, no_alloc=True), got_small)


def test_noalloc_rejects_inputs_beyond_its_scratch(hip_device):
    huge = np.ones((400, 200_000), np.uint8)
    with pytest.raises(B3hError) as e:
        deduplicate(huge, device=0, no_alloc=True)
    assert e.value.status == 1  # B3H_ERR_ARG
    assert "scratch" in str(e.value)
