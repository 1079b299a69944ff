"""GPU parity of the fused star join (join + fetch + 2-key group-by + grouped sum) vs the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen, workloads as W

pytestmark = pytest.mark.gpu

KEYS = ["sums", "groid", "key1", "key2", "e_fk", "e_dk", "e_a1", "e_a2", "e_m", "e_grp"]


def run_both(spec, **kw):
    h = W.star1_host(spec)
    oc = (O.Col(h["fk"], spec.meta_key), O.Col(h["m"], spec.meta_m), O.Col(h["key"], spec.meta_key),
          O.Col(h["a1"], spec.meta_attr), O.Col(h["a2"], spec.meta_attr))
    ref = O.star1_pipeline(*oc)
    dc = (capi.Column.from_numpy(h["fk"], spec.meta_key), capi.Column.from_numpy(h["m"], spec.meta_m),
          capi.Column.from_numpy(h["key"], spec.meta_key), capi.Column.from_numpy(h["a1"], spec.meta_attr),
          capi.Column.from_numpy(h["a2"], spec.meta_attr))
    got = capi.star1_pipeline(*dc, **kw)
    return h, ref, got


@pytest.mark.parametrize("n_fact,n_dim,rate,seed", [
    (0, 10, 0.0, 1), (1, 10, 0.0, 2), (2047, 300, 0.0, 3), (2049, 300, 1e-3, 4), (16_385, 1000, 1e-3, 5),
    (250_000, 5_000, 2e-4, 6), (1_000_003, 200_000, 1e-5, 7), (600_000, 20, 1e-3, 8),
])
def test_star1_parity(n_fact, n_dim, rate, seed):
    spec = W.Star1Spec(n_fact=n_fact, n_dim=n_dim, rate=rate, seed=seed)
    _, ref, got = run_both(spec)
    assert got["status"] == 0
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("card1,card2,n_dim,rate,seed", [
    (18, 7, 3_000, 2e-3, 41),     # 126 combinations: the most the LDS-map probe takes
    (127, 1, 3_000, 2e-3, 42),    # 127 combinations: the general probe
    (25, 5, 300_000, 1e-3, 43),   # keys beyond the LDS-resident part of the map (byte map in HBM)
    (25, 5, 3_000, 1e-2, 44),     # errors in almost every wave: every wave rescanned for positions
])
def test_star1_parity_fast_probe_boundaries(card1, card2, n_dim, rate, seed):
    spec = W.Star1Spec(n_fact=1_000_003, n_dim=n_dim, card1=card1, card2=card2, rate=rate, seed=seed)
    _, ref, got = run_both(spec)
    assert got["status"] == 0
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k


def _offset_column(arr, meta, offset):
    """A device column whose data starts `offset` bytes into its allocation."""
    base = capi.DeviceBuffer.from_numpy(np.concatenate([np.zeros(offset, np.uint8), arr.view(np.uint8)]))
    b = capi.DeviceBuffer.__new__(capi.DeviceBuffer)
    b.t = base.t[offset:]
    b.nbytes = arr.nbytes
    return capi.Column(b, arr.size, meta)


def test_star1_parity_general_probe_on_unaligned_fact_columns():
    # fact columns 8 bytes off a 16-byte boundary: the general probe runs on the same data
    spec = W.Star1Spec(n_fact=700_001, n_dim=5_000, rate=1e-3, seed=45)
    h = W.star1_host(spec)
    oc = (O.Col(h["fk"], spec.meta_key), O.Col(h["m"], spec.meta_m), O.Col(h["key"], spec.meta_key),
          O.Col(h["a1"], spec.meta_attr), O.Col(h["a2"], spec.meta_attr))
    ref = O.star1_pipeline(*oc)
    dc = (_offset_column(h["fk"], spec.meta_key, 8), _offset_column(h["m"], spec.meta_m, 8),
          capi.Column.from_numpy(h["key"], spec.meta_key), capi.Column.from_numpy(h["a1"], spec.meta_attr),
          capi.Column.from_numpy(h["a2"], spec.meta_attr))
    assert dc[0].buf.ptr % 16 == 8
    got = capi.star1_pipeline(*dc)
    assert got["status"] == 0
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k


def test_star1_parity_small_tiles_and_many_groups():
    # 256 x 256 attribute values: up to 65536 groups (more than one tile's LDS table holds)
    spec = W.Star1Spec(n_fact=400_000, n_dim=70_000, card1=256, card2=256, rate=1e-4, seed=11)
    with capi.tiling(2, 8):
        _, ref, got = run_both(spec, groups_cap=16)
    if len(ref["sums"]) > 8192:
        assert got["status"] == 2  # too many groups for the fused finalize: caller falls back
    else:
        assert got["status"] == 0
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), k


def test_star1_duplicate_build_keys_reports_not_applicable():
    spec = W.Star1Spec(n_fact=10_000, n_dim=100, rate=0.0, seed=12)
    h = W.star1_host(spec)
    h["key"][7] = h["key"][3]
    dc = (capi.Column.from_numpy(h["fk"], spec.meta_key), capi.Column.from_numpy(h["m"], spec.meta_m),
          capi.Column.from_numpy(h["key"], spec.meta_key), capi.Column.from_numpy(h["a1"], spec.meta_attr),
          capi.Column.from_numpy(h["a2"], spec.meta_attr))
    got = capi.star1_pipeline(*dc)
    assert got["status"] == 1
    # the next call on the same workspace is unaffected
    spec2 = W.Star1Spec(n_fact=50_000, n_dim=500, rate=1e-3, seed=13)
    _, ref, got2 = run_both(spec2)
    assert got2["status"] == 0
    for k in KEYS:
        assert np.array_equal(got2[k], ref[k]), k


def test_star1_device_generator_matches_host():
    spec = W.Star1Spec(n_fact=300_001, n_dim=4_000, rate=1e-3, seed=21)
    h = W.star1_host(spec)
    dcols = W.star1_device(spec)
    capi.synchronize()
    for name, col in zip(("fk", "m", "key", "a1", "a2"), dcols):
        assert np.array_equal(col.to_numpy(), h[name]), name


def test_star1_repeated_launches_are_identical():
    spec = W.Star1Spec(n_fact=500_000, n_dim=10_000, rate=1e-4, seed=31)
    st = capi.Star1(*W.star1_device(spec))
    st.launch()
    a = st.results()
    for _ in range(3):
        st.launch()
    b = st.results()
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def test_star1_parity_full_size():
    """The benchmark workload itself (60M fact rows, 200k keys, 1e-7 flips), bit-exact vs the oracle."""
    spec = W.Star1Spec()
    print(f"star1 full size: n_fact={spec.n_fact} n_dim={spec.n_dim} rate={spec.rate} seed={spec.seed}", flush=True)
    h, ref, got = run_both(spec)
    assert got["status"] == 0
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert got["sums"].size == spec.card1 * spec.card2
    # every injected flip in the foreign keys is detected (single flips of A=63877 codes, 2^64/A > maxU)
    assert np.array_equal(got["e_fk"], (h["flips"]["fk"].astype(np.uint64) * np.uint64(an.A_OID_ERR)))
