"""Every benchmark configuration at its stated size, through the C-ABI,
against the FP64 oracle (or the golden fixtures the oracle wrote).

  D1  n=20 d=2 chi=32 Gaussian MPS: canonize!(site 1) + overlap + norm
      (canonize.jl:128-149, overlap.jl:36-50, norm.jl:17-37); rel. error
      <= 1e-12 against the oracle, state unchanged, isometry pattern.
  D2  TEBD n=64 depth 20 chi_max=256 threshold 1e-10 (630 gates): relative
      l2 error of the 1024 fixture amplitudes <= 1e-8 and every bond size
      equal to the fixture's (tests/golden, written by the oracle), for the
      FP64 and split-precision contraction forms, and the Vidal-gauge form
      against its own fixture.
  D3  the whole grouped batch (8192 contractions + 8192 truncated SVDs) in
      one call each; sampled problems against the oracle on the same
      (downloaded) complex64 inputs.
  D4  DMRG two-site H_eff at chi=2048 (L, R (2048,5,2048), theta
      (2048,2,2,2048)) in every precision; the first output rows against
      the oracle on the same rows.
  D5  serving states at the configuration's sizes (n 16..128, chi
      16..512): sampled requests against the oracle, and each request
      alone equal to its value in the mix.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import synth
from oracle import tenet_oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BASE = 20261202


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- D1
def test_d1_plain_canonize_overlap_norm(ctx):
    import tenet_hip as T
    from tenet_hip.workloads import plain_state_arrays

    arrs = plain_state_arrays(BASE + 0, 20, 32)
    psi = T.MPS.from_host(ctx, arrs)
    ref = O.OMPS([a.copy() for a in arrs])
    dense0 = ref.to_dense()
    psi.canonize_(0)
    O.canonize_mixed(ref, 0, 0)
    assert psi.form == ("mixed", 0, 0)
    # thin QR shrinks the right-edge bonds to 2, 4, 8, 16 (canonize.jl:58)
    assert psi.bondsizes() == ref.bondsizes()
    assert psi.bondsizes()[-4:] == [16, 8, 4, 2]
    sites = psi.to_host()
    for a in sites[1:]:
        assert O.is_right_isometry(a, atol=1e-12)
    assert _rel(O.OMPS(sites).to_dense(), dense0) < 1e-12
    ov = psi.overlap(psi)
    ov_ref = O.overlap_mps(ref, ref)
    assert abs(ov - ov_ref) <= 1e-12 * abs(ov_ref)
    nr = psi.norm()
    assert abs(nr - O.norm(ref)) <= 1e-12 * O.norm(ref)
    # norm(psi)^2 = overlap(psi, psi) (the invariant norm.jl relies on)
    assert abs(nr * nr - ov.real) <= 1e-12 * abs(ov)


# ---------------------------------------------------------------- D2
def _tebd_inputs():
    from tenet_hip.tebd import brickwork_schedule, haar_gates

    sched = brickwork_schedule(64, 20)
    return sched, haar_gates(BASE + 1, len(sched))


@pytest.mark.parametrize("form,prec", [("fp64", 0), ("split_bf16", 2), ("split_fp16", 3)])
def test_d2_tebd_full_size_vs_fixture(ctx, form, prec):
    from tenet_hip.tebd import tebd_sweep, upload_gates

    fx = np.load(os.path.join(GOLDEN, "tebd_n64_d20_chi256.npz"))
    bits = np.unpackbits(fx["bits"], axis=1)[:, :64]
    sched, gates = _tebd_inputs()
    psi = tebd_sweep(ctx, 64, sched, upload_gates(ctx, gates), 256, 1e-10, prec=prec)
    assert list(psi.bondsizes()) == list(fx["bonds"])
    amps = psi.amplitudes(bits)
    err = _rel(amps, fx["amps"])
    print(f"D2 {form}: amplitude relerr vs FP64 oracle {err:.3e}")
    assert err <= 1e-8


def test_d2_tebd_vidal_full_size_vs_fixture(ctx):
    from tenet_hip.tebd import upload_gates
    from tenet_hip.vidal import brickwork_layers, tebd_vidal_sweep

    fx = np.load(os.path.join(GOLDEN, "tebd_vidal_n64_d20_chi256.npz"))
    bits = np.unpackbits(fx["bits"], axis=1)[:, :64]
    sched, gates = _tebd_inputs()
    # the serpentine schedule's layers, in application order
    layers, i = [], 0
    for par in range(20):
        cnt = len(range(par % 2, 63, 2))
        layers.append(list(sched[i:i + cnt]))
        i += cnt
    assert [sorted(x) for x in layers] == [sorted(x) for x in brickwork_layers(64, 20)]
    v = tebd_vidal_sweep(ctx, 64, layers, upload_gates(ctx, gates), 256, 1e-10)
    assert list(v.bondsizes()) == list(fx["bonds"])
    amps = v.to_mps().amplitudes(bits)
    err = _rel(amps, fx["amps"])
    print(f"D2 Vidal: amplitude relerr vs the Vidal-gauge FP64 oracle {err:.3e}")
    assert err <= 1e-8
    mid = len(v.lam_host) // 2
    assert np.allclose(v.lam_host[mid], fx["lam_mid"][:len(v.lam_host[mid])], rtol=1e-9, atol=1e-12)
    v.close()


# ---------------------------------------------------------------- D3
def test_d3_grouped_batch_full_size(ctx):
    import tenet_hip as T
    from tenet_hip import _lib as L
    from tenet_hip.grouped import GroupedPlan, grouped_svd
    from tenet_hip.workloads import grouped_contractions, grouped_svds

    seed, count = BASE + 2, 8192
    specs = grouped_contractions(seed, count)
    sizes = [sp.sizes for sp in specs]
    in_off, out_off, ai, ao = [], [], 0, 0
    for a, b, c in sizes:
        in_off.append((ai, ai + a))
        ai += a + b
        out_off.append(ao)
        ao += c
    arena_in = T.DeviceArray(ctx, (ai,), np.complex64).gaussian_(seed)
    arena_out = T.DeviceArray(ctx, (ao,), np.complex64)
    plan = GroupedPlan(ctx, [(sp.shape(sp.la), sp.la, sp.shape(sp.lb), sp.lb, sp.shape(sp.lc), sp.lc)
                             for sp in specs], np.complex64)
    ptrs = []
    for (oa, ob), oc in zip(in_off, out_off):
        ptrs += [arena_in.ptr + 8 * oa, arena_in.ptr + 8 * ob, arena_out.ptr + 8 * oc]
    plan.exec(ptrs)

    def dl(ptr, count):
        h = np.empty(count, dtype=np.complex64)
        L.check(L.lib.tenet_memcpy_d2h(ctx.h, h.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), h.nbytes, None))
        return h

    rng = np.random.default_rng(3)
    worst = 0.0
    for i in rng.choice(count, 32, replace=False):
        sp = specs[int(i)]
        a, b, c = sizes[int(i)]
        oa, ob = in_off[int(i)]
        A = dl(arena_in.ptr + 8 * oa, a).reshape(sp.shape(sp.la), order="F").astype(np.complex128)
        B = dl(arena_in.ptr + 8 * ob, b).reshape(sp.shape(sp.lb), order="F").astype(np.complex128)
        ref = O.binary_einsum(O.OTensor(A, sp.la), O.OTensor(B, sp.lb), blas=True).permute(sp.lc).data
        got = dl(arena_out.ptr + 8 * out_off[int(i)], c).reshape(sp.shape(sp.lc), order="F")
        worst = max(worst, _rel(got, ref))
    print(f"D3 contractions: max relerr over 32 sampled problems {worst:.2e}")
    assert worst < 2e-5  # FP32 MFMA on complex64 operands
    plan.close()
    del arena_out

    ms, ns, maxdims = grouped_svds(seed + 1000, count)
    svd_in = T.DeviceArray(ctx, (int(np.sum(ms * ns)),), np.complex64).gaussian_(seed + 1000)
    mats, o = [], 0
    for m, n in zip(ms, ns):
        v = T.DeviceArray(ctx, (int(m), int(n)), np.complex64, ptr=svd_in.ptr + 8 * o, owner=False)
        v._base = svd_in
        mats.append(v)
        o += int(m * n)
    res = grouped_svd(ctx, mats, maxdims, 1e-10)
    worst_s = 0.0
    for i in rng.choice(count, 16, replace=False):
        m, n = int(ms[i]), int(ns[i])
        Ah = mats[int(i)].to_host().astype(np.complex128)
        u, s_ref, vh = np.linalg.svd(Ah, full_matrices=False)
        U, S, Vh, k, disc = res[int(i)]
        assert k == min(int(maxdims[i]), int(np.sum(s_ref > 1e-10)))
        worst_s = max(worst_s, float(np.max(np.abs(S[:k] - s_ref[:k])) / s_ref[0]))
        assert disc == pytest.approx(float(np.sum(s_ref[k:] ** 2)), rel=1e-9)
        best = (u[:, :k] * s_ref[:k]) @ vh[:k]
        assert _rel((U.to_host()[:, :k] * S[:k]) @ Vh.to_host()[:k], best) < 1e-6
    print(f"D3 SVDs: max |sigma - LAPACK| / sigma_1 over 16 sampled problems {worst_s:.2e}")
    assert worst_s < 1e-12


# ---------------------------------------------------------------- D4
@pytest.mark.parametrize("form,prec,tol", [("fp64", 0, 1e-12), ("fp32", 1, 1e-5), ("split_bf16", 2, 1e-10),
                                           ("split_fp16", 3, 1e-10), ("split_i8", 4, 1e-12),
                                           ("split_i8_fast", 5, 5e-9)])
def test_d4_heff_chi2048_vs_oracle_rows(ctx, heff2048, form, prec, tol):
    from tenet_hip.dmrg import heff_apply

    tL, tW1, tW2, tR, tth, ref_rows = heff2048
    y = heff_apply(tL, tW1, tW2, tR, tth, prec=prec)
    yh = y.to_host()[:ref_rows.shape[0]]
    err = _rel(yh, ref_rows)
    print(f"D4 {form}: relerr of the first {ref_rows.shape[0]} output rows {err:.3e}")
    assert err < tol


@pytest.fixture(scope="module")
def heff2048(ctx):
    import tenet_hip as T
    from tenet_hip.dmrg import xxz_mpo_tensor

    chi, w, seed, rows = 2048, 5, BASE + 3, 32
    Ld = T.DeviceArray(ctx, (chi, w, chi)).gaussian_(seed)
    Rd = T.DeviceArray(ctx, (chi, w, chi)).gaussian_(seed + 1)
    th = T.DeviceArray(ctx, (chi, 2, 2, chi)).gaussian_(seed + 2)
    nrm = np.sqrt(th.nrm2sq())
    th.scale_(1.0 / nrm)
    Wh = xxz_mpo_tensor(1.0, 0.5)
    tL = T.Tensor(Ld, ["l", "w", "lk"])
    tR = T.Tensor(Rd, ["r", "wr", "rk"])
    tth = T.Tensor(th, ["lk", "i1", "i2", "rk"])
    tW1 = T.Tensor.from_host(ctx, Wh, ["w", "m", "o1", "i1"])
    tW2 = T.Tensor.from_host(ctx, Wh, ["m", "wr", "o2", "i2"])
    # the oracle on the same (host-regenerated) inputs, first `rows` output rows
    Lh = synth.gaussian(seed, chi * w * chi).reshape(chi, w, chi, order="F")[:rows].copy()
    Rh = synth.gaussian(seed + 1, chi * w * chi).reshape(chi, w, chi, order="F")
    thh = synth.gaussian(seed + 2, chi * 4 * chi).reshape(chi, 2, 2, chi, order="F") / nrm
    ref_rows = O.heff_apply(Lh, Wh, Wh, Rh, thh)
    return tL, tW1, tW2, tR, tth, ref_rows


# ---------------------------------------------------------------- D5
def test_d5_serving_full_size_states(ctx):
    from tenet_hip.serving import Server
    from tenet_hip.workloads import build_serving_states, serving_requests

    # six states drawn as the serving workload draws them, plus one at the
    # largest stated size (n = 128, chi = 512: 0.54 GB in complex64)
    states = build_serving_states(ctx, BASE + 4, 6)
    states += build_serving_states(ctx, BASE + 104, 1, n_lo=128, n_hi=128, chi_lo=512, chi_hi=512)
    chis = [int(max(s.chis)) for s in states]
    ns = [s.n for s in states]
    print(f"D5 states: n {ns}, chi {chis}, {sum(s.bytes for s in states) / 1e9:.2f} GB")
    assert min(ns) >= 16 and max(ns) == 128 and max(chis) == 512
    reqs, arr = serving_requests(BASE + 5, states, 400, rate=4000.0)
    srv = Server(ctx, states, n_streams=4, max_inflight=256, max_amp_batch=128)
    srv.run(reqs, arr)
    srv.close()
    hosts = {}
    rng = np.random.default_rng(8)
    amp_idx = [i for i, r in enumerate(reqs) if r.kind == "amp"]
    pau_idx = [i for i, r in enumerate(reqs) if r.kind == "pauli"]
    big = len(states) - 1
    pick = list(rng.choice(amp_idx, 10, replace=False)) + list(rng.choice(pau_idx, 12, replace=False))
    pick += [i for i in amp_idx if reqs[i].state == big and i not in pick][:2]
    pick += [i for i in pau_idx if reqs[i].state == big and i not in pick][:2]
    # oracle sweeps cost ~4 n chi^3 complex MACs: expectations are checked on the
    # states whose sweep stays under ~2e10 MACs (every amplitude is checked)
    cost = [4.0 * float(np.sum(s.chis.astype(np.float64) ** 3)) for s in states]
    alone = Server(ctx, states, n_streams=1)
    checked = {"amp": 0, "pauli": 0}
    for i in pick:
        r = reqs[int(i)]
        if r.kind == "amp" or cost[r.state] <= 2e10:
            if r.state not in hosts:
                hosts[r.state] = O.OMPS([a.to_host().astype(np.complex128) for a in states[r.state].mps.sites])
            psi = hosts[r.state]
            ref = O.amplitude(r.bits, psi) if r.kind == "amp" else O.expectation_pauli(psi, r.start, r.paulis)
            tol = 1e-4 * max(abs(ref), 1e-30) + (1e-6 if r.kind == "pauli" else 0.0)
            assert abs(r.result - ref) <= tol, (r.kind, r.rid, r.state, r.result, ref)
            checked[r.kind] += 1
        # the same request alone in a fresh server equals its value in the mix
        solo = type(r)(r.rid, r.state, r.kind, bits=r.bits, start=r.start, paulis=r.paulis)
        alone.run([solo])
        assert abs(solo.result - r.result) <= 1e-6 * max(abs(r.result), 1e-30), (r.kind, r.rid)
    print(f"D5: oracle-checked {checked}, alone-vs-mix {len(pick)}")
    n_amp = sum(1 for i in pick if reqs[int(i)].kind == "amp")
    assert checked["amp"] == n_amp >= 12 and checked["pauli"] >= 3
    assert any(reqs[int(i)].state == big and reqs[int(i)].kind == "amp" for i in pick)
    alone.close()
    for st in states:
        st.close()
