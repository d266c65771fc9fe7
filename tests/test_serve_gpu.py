"""GPU tests of the serving loop (loops_server_*): every result is the same
alone and in the mix (bit-identical), meets the oracle's 1e-5 bar against
FP64 (fp32 and bf16 / fp16 storage, stored values), equals the per-call
plain mode bit for bit, and requests complete under memory admission with a
small arena and under an open-loop replay."""
import numpy as np
import pytest

import oracle
from conftest import random_csr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import edge_loops_amd as loops  # noqa: E402
from edge_loops_amd import synth  # noqa: E402
from edge_loops_amd.serve import Server, poisson_arrivals  # noqa: E402


def _pool():
    mats = []
    for i in range(4):
        off, idx, val = random_csr(3000 + 2000 * i, 4000 + 1000 * i, 20 + 10 * i, seed=70 + i,
                                   long_rows=((i, 3500),))
        mats.append((off, idx, val, 4000 + 1000 * i))
    B = synth.powerlaw_csr(60000, 60000, 2.0, 1, 20000, 600000, seed=71)
    mats.append((B.offsets, B.indices, B.values, B.cols))
    mats.append((np.zeros(51, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 7))
    return mats


MATS = _pool()
KS = [1, 3, 16, 1, 32, 5, 1, 64, 2, 1, 17, 8]


def _setup(srv, vt):
    rng = np.random.default_rng(5)
    ids, dev = [], []
    for off, idx, val, cols in MATS:
        A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val, value_type=vt)
        ids.append(srv.add_matrix(A))
        dev.append(A)
    reqs = []
    for j, k in enumerate(KS * 2):
        m = j % len(MATS)
        cols = MATS[m][3]
        xh = rng.uniform(-1, 1, cols if k == 1 else (cols, k)).astype(np.float32)
        reqs.append((ids[m], m, torch.from_numpy(xh).cuda(), xh))
    return dev, reqs


def _check_oracle(dev, m, xh, y):
    off, idx, _, _ = MATS[m]
    stored = dev[m].values.float().cpu().numpy()
    if xh.ndim == 1:
        r, a = oracle.spmv_f64(off, idx, stored, xh)
    else:
        r, a = oracle.spmm_f64(off, idx, stored, xh)
    assert np.all(np.abs(y - r) <= 1e-5 * a + 1e-30)


@pytest.mark.parametrize("vt", ["f32", "bf16", "f16"])
def test_same_alone_and_in_the_mix(vt):
    with Server() as srv:
        dev, reqs = _setup(srv, vt)
        alone = []
        for mid, m, x, xh in reqs:
            alone.append(srv.result(srv.submit(mid, x)).cpu().numpy())
        tickets = [srv.submit(mid, x) for mid, m, x, xh in reqs]
        mixed = [srv.result(t).cpu().numpy() for t in tickets]
        st = srv.stats()
        assert st["completed"] == 2 * len(reqs)
        assert st["batches"] < st["completed"]  # the burst was coalesced
        for (mid, m, x, xh), a, b in zip(reqs, alone, mixed):
            assert np.array_equal(a, b), (m, xh.shape)
            _check_oracle(dev, m, xh, b)


def test_plain_mode_equals_batched():
    # batched mode launches the large requests alone and batches the rest;
    # batch_all puts every request in the batch: all three bit-identical
    outs = {}
    for name, mode, batch_all in (("batched", "batched", False), ("plain", "plain", False),
                                  ("batch_all", "batched", True)):
        with Server(mode=mode, batch_all=batch_all) as srv:
            dev, reqs = _setup(srv, "f32")
            tickets = [srv.submit(mid, x) for mid, m, x, xh in reqs]
            outs[name] = [srv.result(t).cpu().numpy() for t in tickets]
            if mode == "plain":
                st = srv.stats()
                assert st["batches"] == st["completed"] == len(reqs)
    for a, b, c in zip(outs["batched"], outs["plain"], outs["batch_all"]):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    with pytest.raises(loops.LoopsError):
        Server(mode="plain", batch_all=True)


def test_plain_reference_schedules():
    # Plain mode through the reference's own entry points: k == 1 via
    # merge_path_flat (default tiles), k > 1 via thread_mapped SpMM, which
    # sums each row in CSR order (bit-identical to the fp32 oracle).
    with Server(mode="plain", reference_schedules=True) as srv:
        dev, reqs = _setup(srv, "f32")
        tickets = [srv.submit(mid, x) for mid, m, x, xh in reqs]
        outs = [srv.result(t).cpu().numpy() for t in tickets]
    for (mid, m, x, xh), y in zip(reqs, outs):
        _check_oracle(dev, m, xh, y)
        if xh.ndim == 2:
            off, idx, val, _ = MATS[m]
            assert np.array_equal(y, oracle.spmm_f32(off, idx, val, xh))
    with pytest.raises(loops.LoopsError):
        Server(mode="batched", reference_schedules=True)


def test_memory_admission_small_arena():
    biggest = max((off.shape[0] - 1) * k for off, *_ in MATS for k in KS) * 4
    arena = 2 * ((biggest + 255) // 256 * 256)
    with Server(arena_bytes=arena, slots=3) as srv:
        dev, reqs = _setup(srv, "f32")
        tickets = [srv.submit(mid, x) for mid, m, x, xh in reqs]
        for t, (mid, m, x, xh) in zip(tickets, reqs):
            y = srv.result(t).cpu().numpy()
            _check_oracle(dev, m, xh, y)
        st = srv.stats()
        assert st["completed"] == len(reqs)
        assert st["arena_high_water"] <= arena


def test_replay_open_loop():
    with Server(slots=2) as srv:
        dev, reqs = _setup(srv, "f32")
        load = [(mid, x) for mid, m, x, xh in reqs] * 4
        arrivals = poisson_arrivals(len(load), 20000.0, seed=9)
        lat, span = srv.replay(load, arrivals)
        assert lat.shape == (len(load),) and np.all(lat > 0) and span > 0
        st = srv.stats()
        assert st["completed"] == st["submitted"] == len(load)


def test_rejects_bad_requests():
    with Server(arena_bytes=1 << 20) as srv:
        off, idx, val, cols = MATS[4]
        A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
        mid = srv.add_matrix(A)
        x = torch.zeros(cols, 64, device="cuda")
        with pytest.raises(loops.LoopsError):   # output never fits the arena
            srv.submit(mid, x)
        with pytest.raises(loops.LoopsError):
            srv.submit(mid + 5, x[:, 0].contiguous())


def test_input_still_in_production():
    """x is written by work still queued on another stream when the request
    is submitted: the event submit() records there orders the server's read
    after it, so the result is the product with the final x."""
    off, idx, val, cols = MATS[4]
    A = loops.csr_t.from_host(off.shape[0] - 1, cols, off, idx, val)
    xh = np.random.default_rng(17).uniform(-1, 1, (cols, 4)).astype(np.float32)
    xd = torch.from_numpy(xh).cuda()
    torch.cuda.synchronize()
    ref = oracle.spmm_f64(off, idx, val, xh)
    with Server() as srv:
        mid = srv.add_matrix(A)
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            x = torch.zeros((cols, 4), device="cuda")
            busy = torch.full((4096, 4096), 1e-3, device="cuda")
            for _ in range(12):  # milliseconds of queued work ahead of the copy
                busy = busy @ busy
            x.copy_(xd)
        t = srv.submit(mid, x, stream=side)
        y = srv.result(t).cpu().numpy()
        side.synchronize()
    assert np.all(np.abs(y - ref[0]) <= 1e-5 * ref[1] + 1e-30)
