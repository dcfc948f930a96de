"""The autotuner and the BSGS shapes it picks from (include/mps_gpu.h mps_gpu_autotune,
mps_gpu_set_tuning): every shape, under class policy 1 (class mode) and 2 (order mode), gives the
oracle's relations, the tuner installs the fastest for the item's key and only for that key, and
invalid shapes are rejected."""
import pytest

pytestmark = pytest.mark.gpu

CLASS_ITEM = (10, 111, 10**12, 10**12 + 2**21, 0, 10_000_000, 6, 4)     # class mode
PLAIN_ITEM = (2, -25, 10**9, 10**9 + 2**21, 3_600_000, 5_000_000, 1, 0)  # class mode under policy 1
SHORT_ITEM = (3, 5, 10**10, 10**10 + 2**21, 1000, 50_000, 1, 0)          # span too short: plain mode


def oracle_rels(oracle, it):
    b, c, P0, P1, N0, N1, mod, res = it
    return oracle.process_range(oracle.make_item(b, c, P0, P1, N0, N1, mod, res), P0, P1, 1 << 20, 8)[0]


@pytest.fixture(autouse=True)
def clean_tuning(gpu):
    gpu.clear_tuning()
    yield
    gpu.clear_tuning()


@pytest.mark.parametrize("it", [CLASS_ITEM, PLAIN_ITEM, SHORT_ITEM])
@pytest.mark.parametrize("shape", [dict(waves=1, load=4, cmax=512, mcap=1024, policy=1, ext=1024),
                                   dict(waves=4, load=2, cmax=2048, mcap=2048, policy=2, ext=1),
                                   dict(waves=2, load=4, cmax=1024, mcap=4096, policy=1, ext=1024),
                                   dict(waves=1, load=2, cmax=512, mcap=1024, policy=2, ext=1669),
                                   dict(waves=2, load=2, cmax=1024, mcap=512, policy=2, ext=100)])
def test_every_shape_gives_oracle_relations(gpu, mps, oracle, it, shape):
    """Every BSGS shape, class policy and medium-prime bound (ext: 1 = none, up to the whole
    table) gives the oracle's relations."""
    print(f"shape {shape} item {it}", flush=True)
    w = mps.work_item(*it)
    gpu.set_tuning(w, **shape)
    assert gpu.get_tuning(w)[0] == shape
    rels, st = gpu.run_item(w, chunk_ints=1 << 21)
    assert st.declined == 0
    assert rels == oracle_rels(oracle, it)


def test_autotune_installs_fastest_for_key(gpu, mps, oracle):
    w = mps.work_item(*CLASS_ITEM)
    other = mps.work_item(*SHORT_ITEM)
    _, key = gpu.get_tuning(w)
    _, key_other = gpu.get_tuning(other)
    assert key != key_other
    before_other = gpu.get_tuning(other)[0]
    best, best_ms, default_ms = gpu.autotune(w, sample_odds=1 << 20)
    print(f"autotune {best} {best_ms:.3f} ms (default {default_ms:.3f} ms)", flush=True)
    assert 0 < best_ms <= default_ms
    assert gpu.get_tuning(w) == (best, key)
    assert gpu.get_tuning(other)[0] == before_other
    rels, _ = gpu.run_item(w, chunk_ints=1 << 21)
    assert rels == oracle_rels(oracle, CLASS_ITEM)


def test_invalid_shapes_rejected(gpu, mps):
    w = mps.work_item(*CLASS_ITEM)
    for bad in (dict(waves=3), dict(load=3), dict(cmax=4096), dict(mcap=100), dict(policy=3)):
        with pytest.raises(ValueError):
            gpu.set_tuning(w, **bad)


def test_invalid_configuration_rejected_whole(gpu, mps, oracle):
    """mps_gpu_configure checks every argument before any takes effect: a call with one bad
    argument changes nothing (here the Montgomery field a good field argument would force), and the
    handle then still gives the oracle's relations."""
    for bad in (dict(field=2), dict(field=-2), dict(field=0, waves=3), dict(field=0, mcap=100),
                dict(field=0, pass_odds=100), dict(field=0, pass_odds=1 << 32)):
        assert gpu.configure(**bad) == mps.MPS_ERROR, bad
    P0, P1 = 10**12, 10**12 + 2**20
    item = (10, 111, P0, P1, 0, 10_000_000, 6, 4)
    gpu.reset_stats()
    rels, _ = gpu.run_item(mps.work_item(*item), chunk_ints=1 << 20)
    assert gpu.stats()["field"] == 1  # still the automatic choice: FP64 below 2^50
    want, _ = oracle.process_range(oracle.make_item(*item), P0, P1, 1 << 20, 8)
    assert rels == want
