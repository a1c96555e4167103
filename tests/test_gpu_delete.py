"""rtrex_delete_and_clean (gpuDeleteAndClean / deleteAndClean,
Decomposition.h:147-283) on the GPU against the oracle: several calls in a
row, pairs given in either orientation (the reference's "lower" endpoint on a
degree tie is the pair's second), then the extraction from that state.

Every call is compared with the oracle's state after the same calls.  In the
checking mode RTREX_REPLAY_ALL=1 every region a deleted edge belongs to is
re-run in the reference's own order (the initial clean, then each logged
call's pushes and LIFO cascade), so the replay of logged deletions is itself
exercised on every call; without it the certified decisions run, and any
undecided one takes the same replay.  Bad pair lists are refused before the
graph changes."""
from __future__ import annotations

import zlib

import numpy as np
import pytest

import oracle as O
import rtrex_cuda_amd as R
from conftest import golden_graph

pytestmark = pytest.mark.gpu

CASES = [("clean_email", 0.1), ("clean_astro", 0.1), ("clean_email", 0.3)]


def _calls(off, nb, eps, state, seed, n_calls=3, per_call=25):
    """Deletion calls: alive edges picked with a seeded generator, each pair in
    a random orientation; later calls pick among edges alive in the oracle's
    state after the earlier ones."""
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    out, st = [], state
    for _ in range(n_calls):
        alive = np.nonzero((st.status == ord("Y")) & (nb > src))[0]
        pick = rng.choice(alive, size=min(per_call, len(alive)), replace=False)
        pairs = np.stack([src[pick], nb[pick]], 1)
        flip = rng.random(len(pairs)) < 0.5
        pairs[flip] = pairs[flip][:, ::-1]
        st = O.delete_and_clean(off, nb, eps, pairs, st)
        out.append((pairs, st))
    return out


@pytest.mark.parametrize("replay_all", [0, 1])
@pytest.mark.parametrize("name,eps", CASES)
def test_delete_calls_match_oracle_then_extract(name, eps, replay_all, monkeypatch):
    monkeypatch.setenv("RTREX_REPLAY_ALL", str(replay_all))
    n, u, v = golden_graph(name)
    off, nb = O.csr_from_undirected(n, u, v)
    cl = O.initial_clean(off, nb, eps)
    calls = _calls(off, nb, eps, cl, seed=zlib.crc32(f"{name}{eps}".encode()))
    print(f"[gpu] delete {name} eps {eps} replay_all {replay_all}", flush=True)
    with R.DeviceGraph.from_csr(off, nb) as g:
        g.gpu_common_nbr()
        g.gpu_initial_clean(eps)
        for i, (pairs, exp) in enumerate(calls):
            g.gpu_delete_and_clean(pairs, eps)
            st = g.sync_edge_status()
            assert np.array_equal(st == ord("Y"), exp.status == ord("Y")), f"call {i}"
            assert g.audit()["ambiguous_clean"] == 0
        cm, k = g.disjoint_extract(eps)
        a = g.audit()
    cm0, k0 = O.extract_from_state(off, nb, eps, calls[-1][1])
    assert k == k0 and np.array_equal(cm, cm0)
    print(f"[gpu]   audit {a}", flush=True)


def test_bad_pairs_refused_before_any_change():
    n, u, v = golden_graph("clean_astro")
    off, nb = O.csr_from_undirected(n, u, v)
    eps = 0.3
    cl = O.initial_clean(off, nb, eps)
    src = np.repeat(np.arange(n), np.diff(off))
    alive = np.nonzero((cl.status == ord("Y")) & (nb > src))[0]
    dead = np.nonzero((cl.status == ord("N")) & (nb > src))[0]
    good = np.stack([src[alive[:5]], nb[alive[:5]]], 1)
    with R.DeviceGraph.from_csr(off, nb) as g:
        g.gpu_common_nbr()
        g.gpu_initial_clean(eps)
        before = g.sync_edge_status()
        bad_lists = [
            np.concatenate([good, good[:1]]),                         # listed twice
            np.concatenate([good, [[src[dead[0]], nb[dead[0]]]]]),    # not alive
            np.concatenate([good, [[0, 0]]]),                         # not an edge
        ]
        for bad in bad_lists:
            with pytest.raises(R.RtrexError) as e:
                g.gpu_delete_and_clean(bad, eps)
            assert e.value.status == 1
            assert np.array_equal(g.sync_edge_status(), before)
        # the handle still works: the good pairs delete as the oracle does
        g.gpu_delete_and_clean(good, eps)
        exp = O.delete_and_clean(off, nb, eps, good, cl)
        assert np.array_equal(g.sync_edge_status() == ord("Y"), exp.status == ord("Y"))
