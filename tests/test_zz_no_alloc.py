"""hj_directory_probe_noalloc (the variant of hj_directory_probe that allocates nothing, its
staging in a fixed device buffer) against hj_directory_probe, through the C ABI: two calls in a
row, the second smaller, then the largest accepted call, and the refusal beyond it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(hjo):
    c = hjo.Context(arena_bytes=1 << 30)
    yield c
    c.close()


def test_noalloc_probe_matches_probe(ctx, hjo):
    rng = np.random.default_rng(77)
    keys = rng.integers(0, 150_000, 200_000).astype(np.uint32)  # repeated keys
    d = hjo.Directory(ctx, keys)
    try:
        for m in (1_000_000, 10_000):  # the second call smaller than the first
            probe = rng.integers(0, 300_000, m).astype(np.uint32)
            want_c, want_f = d.probe(probe)
            got_c, got_f = d.probe_noalloc(probe)
            assert np.array_equal(got_c, want_c)
            assert np.array_equal(got_f, want_f)
            assert 0 < int((got_c > 0).sum()) < m
        big = rng.integers(0, 300_000, hjo.abi.NOALLOC_MAX_PROBE_KEYS).astype(np.uint32)
        want_c, want_f = d.probe(big)
        got_c, got_f = d.probe_noalloc(big)
        assert np.array_equal(got_c, want_c) and np.array_equal(got_f, want_f)
        with pytest.raises(hjo.EngineError) as e:
            d.probe_noalloc(np.zeros(hjo.abi.NOALLOC_MAX_PROBE_KEYS + 1, np.uint32))
        assert e.value.code == hjo.abi.HJ_E_INVALID and "exceed the limit" in e.value.msg
        # the context and the staging stay usable after the refusal
        probe = rng.integers(0, 300_000, 5_000).astype(np.uint32)
        assert np.array_equal(d.probe_noalloc(probe)[0], d.probe(probe)[0])
    finally:
        d.free()
