"""The continuous-batching request loop: every result against the oracle on
the same (downloaded) states, each result the same alone and in the mix, and
admission by HBM footprint.

Tolerance: complex64 states; amplitudes and expectations within 1e-4
relative of the FP64 oracle on the same complex64 data (FP32 accumulation
over <= 64 sites).  Alone vs mix: the kernels compute a request the same way
whatever else is in flight, so the two runs must agree to 1e-6 relative
(bit-identical in practice).
"""
import numpy as np
import pytest

from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(ctx):
    from tenet_hip.workloads import build_serving_states

    states = build_serving_states(ctx, 4242, 6, n_lo=12, n_hi=20, chi_lo=4, chi_hi=24)
    host = [O.OMPS([a.to_host().astype(np.complex128) for a in st.mps.sites]) for st in states]
    return states, host


def _oracle_value(host, r):
    psi = host[r.state]
    if r.kind == "amp":
        return O.amplitude(r.bits, psi)
    return O.expectation_pauli(psi, r.start, r.paulis)


def test_states_are_canonical_and_normalised(served):
    states, host = served
    for psi in host:
        assert abs(O.overlap_mps(psi, psi) - 1.0) < 1e-5
        for a in psi.A[1:]:
            assert O.is_right_isometry(a, atol=1e-5)


def test_serving_results_vs_oracle_and_alone_vs_mix(ctx, served):
    from tenet_hip.serving import Server, latency_summary
    from tenet_hip.workloads import serving_requests

    states, host = served
    reqs, arr = serving_requests(99, states, 300, rate=20000.0)
    srv = Server(ctx, states, n_streams=4, max_inflight=64, max_amp_batch=32)
    srv.run(reqs, arr)
    mix = {r.rid: r.result for r in reqs}
    assert all(r.t_done >= r.t_arrive for r in reqs)
    assert srv.stats["amp_batches"] >= 1 and srv.stats["pauli_launches"] >= 1
    assert srv.stats["env_hits"] >= 1
    for r in reqs:
        ref = _oracle_value(host, r)
        assert abs(r.result - ref) <= 1e-4 * max(abs(ref), 1e-30) + 1e-6 * (r.kind == "pauli"), (r.kind, r.rid)
    srv.close()
    # each request alone, in a fresh server (fresh environment cache)
    alone = Server(ctx, states, n_streams=1)
    for r in reqs[:120]:
        alone.run([r])
        assert abs(r.result - mix[r.rid]) <= 1e-6 * max(abs(mix[r.rid]), 1e-30), r.rid
    alone.close()
    lat = latency_summary(reqs)
    assert lat["p99_ms"] >= lat["p50_ms"] > 0


def test_admission_by_hbm_footprint(ctx, served):
    """A budget just above the states' bytes admits only a few requests at a
    time: the loop waits for completions and still serves everything."""
    from tenet_hip.serving import Server, request_footprint
    from tenet_hip.workloads import serving_requests

    states, host = served
    reqs = serving_requests(7, states, 80)
    state_bytes = sum(s.bytes for s in states)
    env_cap = 1 << 16
    budget = state_bytes + env_cap + 3 * max(request_footprint(states[r.state], r) for r in reqs)
    srv = Server(ctx, states, n_streams=2, hbm_budget=budget, env_cache_bytes=env_cap)
    srv.run(reqs)
    assert srv.stats["admission_waits"] > 0
    for r in reqs:
        ref = _oracle_value(host, r)
        assert abs(r.result - ref) <= 1e-4 * max(abs(ref), 1e-30) + 1e-6 * (r.kind == "pauli")
    srv.close()


def test_prefilled_environments(ctx, served):
    """tenet_left_envs writes every prefix environment in one sweep; a server
    holding them all returns the same values as one that builds them on
    demand."""
    from tenet_hip.serving import Server
    from tenet_hip.workloads import serving_requests

    states, host = served
    reqs = serving_requests(8, states, 120, amp_fraction=0.2)
    lazy = Server(ctx, states, n_streams=3)
    lazy.run(reqs)
    ref = {r.rid: r.result for r in reqs}
    lazy.close()
    full = Server(ctx, states, n_streams=3, prefill_envs=True)
    assert len(full.env_fixed) == sum(st.n - 1 for st in states)
    full.run(reqs)
    assert full.stats["env_built"] == 0
    for r in reqs:
        assert abs(r.result - ref[r.rid]) <= 1e-5 * max(abs(ref[r.rid]), 1e-30) + 1e-7
    full.close()


def test_sample_requests_in_the_mix(ctx, served):
    """Sample requests (sample(psi, nsamples), src/Operations/sample.jl:20-72)
    served together with amplitudes and expectations: each equals MPS.sample
    of the same device state value for value (the same kernels on the same
    data and uniform stream), and the oracle's draws on the downloaded state
    except where a uniform falls within complex64 rounding of a cumulative
    boundary (at most 1% of the rows)."""
    from tenet_hip.serving import Request, Server
    from tenet_hip.workloads import serving_requests

    states, host = served
    reqs, arr = serving_requests(7, states, 60, rate=20000.0)
    samp = [Request(rid=len(reqs) + i, state=i % len(states), kind="sample", nsamples=64 + 32 * i, seed=1000 + i)
            for i in range(6)]
    arrivals = np.concatenate([np.asarray(arr, dtype=float), np.linspace(0.0, float(np.max(arr)), len(samp))])
    srv = Server(ctx, states, n_streams=4, max_inflight=64, max_amp_batch=32)
    srv.run(reqs + samp, arrivals)
    assert srv.stats["sample_launches"] == len(samp)
    assert all(r.t_done >= r.t_arrive for r in reqs + samp)
    for r in samp:
        got = r.result
        assert got.shape == (r.nsamples, states[r.state].n)
        assert np.array_equal(got, states[r.state].mps.sample(r.nsamples, seed=r.seed))
        ora = O.sample(host[r.state], r.nsamples, r.seed)
        assert (ora != got).any(axis=1).mean() <= 0.01
    for r in reqs:
        ref = _oracle_value(host, r)
        assert abs(r.result - ref) <= 1e-4 * max(abs(ref), 1e-30) + 1e-6 * (r.kind == "pauli"), (r.kind, r.rid)
    srv.close()
