"""GPU tests of the serving loop (continuous batching over resident grids).

Each result must be the same when its request is served alone and when it is
served in a mix of concurrent requests: FP32 tile arithmetic into FP64 grids
within 1e-5 relative max-norm of the FP64 oracle (and of each other), native
FP64 arithmetic within 1e-12.
"""
import threading

import numpy as np
import pytest

import migrid
import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(128, 8, 3000), (256, 16, 6000), (192, 32, 4000), (300, 64, 5000), (96, 10, 800),
          (512, 64, 20000)]


def torch():
    import torch as t
    return t


def inputs(i, N, S, n, seed=0):
    if i % 2:
        uv, vals = O.gen_clustered(n, N, S, mean_r=N / 8, seed=seed + i)
    else:
        uv, vals = O.gen_uniform(n, -S / 2, N + S / 2, seed=seed + i)
    return uv, vals, O.config_gcf(S)


def ref(uv, vals, gcf, N, S):
    p = O.grid_zz(uv, vals, N, gcf, S)
    return p[0] + 1j * p[1]


def test_result_is_the_same_alone_and_in_the_mix():
    N, S, n = 256, 32, 8000
    uv, vals, gcf = inputs(1, N, S, n, seed=40)
    want = ref(uv, vals, gcf, N, S)
    with migrid.Server() as srv:
        srv.wait(srv.submit(7, uv, vals, gcf, N, S))
        alone = srv.read_grid(7)
    assert O.rel_errors(alone, want)[0] < 1e-5

    # the same request among concurrent ones from several threads (device
    # submissions, so a burst outpaces the device and shares launches)
    t = torch()
    others = [(100 + i, *SHAPES[i % len(SHAPES)]) for i in range(24)]
    data = {jid: inputs(k, N2, S2, n2, seed=k) + (N2, S2) for k, (jid, N2, S2, n2) in enumerate(others)}
    dev = {jid: [t.from_numpy(np.ascontiguousarray(x)).cuda() for x in d[:3]] for jid, d in data.items()}
    t.cuda.synchronize()
    with migrid.Server(max_batch=8) as srv:
        tickets, lock = [], threading.Lock()

        def client(ids):
            for jid in ids:
                N2, S2 = data[jid][3:]
                tk = srv.submit(jid, *dev[jid], N2, S2)
                with lock:
                    tickets.append(tk)

        ids = list(data)
        th = [threading.Thread(target=client, args=(ids[k::3],)) for k in range(3)]
        for c in th:
            c.start()
        mine = srv.submit(7, uv, vals, gcf, N, S)
        for c in th:
            c.join()
        for tk in tickets + [mine]:
            srv.wait(tk)
        mixed = srv.read_grid(7)
        st = srv.stats()
        assert st["completed"] == len(others) + 1 and st["failed"] == 0
        assert st["batches"] < st["completed"]  # requests shared launches
        for jid, (u, v, g, N2, S2) in list(data.items())[:6]:
            assert O.rel_errors(srv.read_grid(jid), ref(u, v, g, N2, S2))[0] < 1e-5, jid
    assert O.rel_errors(mixed, want)[0] < 1e-5
    assert O.rel_errors(mixed, alone)[0] < 1e-5


def test_grids_accumulate_across_requests_of_a_job():
    N, S = 192, 16
    parts = [inputs(i, N, S, 3000 + 500 * i, seed=9) for i in range(3)]
    want = sum(ref(u, v, g, N, S) for u, v, g in parts)
    with migrid.Server(arith=migrid.ARITH_NATIVE) as srv:
        ts = [srv.submit(3, u, v, g, N, S) for u, v, g in parts]
        for t in ts:
            srv.wait(t)
        got = srv.read_grid(3)
        assert O.rel_errors(got, want)[0] < 1e-12
        srv.drop_job(3)
        assert srv.stats()["jobs_resident"] == 0
        # a dropped job starts again from zero
        u, v, g = parts[0]
        srv.wait(srv.submit(3, u, v, g, N, S))
        assert O.rel_errors(srv.read_grid(3), ref(u, v, g, N, S))[0] < 1e-12


def test_device_and_host_submissions_agree():
    t = torch()
    N, S, n = 300, 64, 5000
    uv, vals, gcf = inputs(3, N, S, n, seed=5)
    want = ref(uv, vals, gcf, N, S)
    with migrid.Server(arith=migrid.ARITH_NATIVE) as srv:
        srv.wait(srv.submit(1, uv, vals, gcf, N, S))
        d = [t.from_numpy(np.ascontiguousarray(x)).cuda() for x in (uv, vals, gcf)]
        t.cuda.synchronize()
        srv.wait(srv.submit(2, *d, N, S))
        a, b = srv.read_grid(1), srv.read_grid(2)
    assert O.rel_errors(a, want)[0] < 1e-12
    assert O.rel_errors(b, want)[0] < 1e-12


def test_admission_waits_for_memory_and_rejects_what_never_fits():
    N, S, n = 1024, 16, 4000
    # room for one 16 MiB grid with its request, not for two
    with migrid.Server(hbm_budget=32 * 2**20, max_inflight=1) as srv:
        parts = [inputs(i, N, S, n, seed=21) for i in range(2)]
        t1 = srv.submit(1, *parts[0], N, S)
        srv.wait(t1)
        # job 2's grid does not fit beside job 1's: it waits until job 1 is dropped
        t2 = srv.submit(2, *parts[1], N, S)
        done, _ = srv.poll(t2)
        assert not done
        assert srv.stats()["queued"] == 1
        srv.drop_job(1)
        srv.wait(t2)
        assert O.rel_errors(srv.read_grid(2), ref(*parts[1], N, S))[0] < 1e-5
        big = inputs(0, 2048, S, 100, seed=1)  # a 64 MiB grid
        t3 = srv.submit(3, *big, 2048, S)
        with pytest.raises(migrid.MigridError) as e:
            srv.wait(t3)
        assert e.value.code == -5
        assert srv.stats()["failed"] == 1


def test_bad_requests_fail_alone():
    N, S = 128, 8
    uv, vals, gcf = inputs(0, N, S, 500, seed=2)
    with migrid.Server() as srv:
        srv.wait(srv.submit(1, uv, vals, gcf, N, S))
        t = srv.submit(1, uv, vals, gcf, 2 * N, S)  # job 1 holds an N x N grid
        with pytest.raises(migrid.MigridError):
            srv.wait(t)
        with pytest.raises(migrid.MigridError):
            srv.submit(1, uv, vals, gcf, N, 7)  # odd support: refused at submit
        srv.wait(srv.submit(1, uv, vals, gcf, N, S))
        assert O.rel_errors(srv.read_grid(1), 2 * ref(uv, vals, gcf, N, S))[0] < 1e-5


def test_completed_statuses_are_bounded():
    """Past 65536 completed requests the oldest statuses are released: an old
    ticket reads as released (an error, not a hang), a recent one as done, and
    every request still completes and adds into its grid."""
    t = torch()
    N, S, jobs, n = 32, 8, 64, 70000
    g_np = O.config_gcf(S)
    uv_np = np.array([[16.25, 15.5]])
    v_np = np.array([[1.0, -0.5]])
    gcf = t.from_numpy(g_np).cuda()
    uv = t.from_numpy(uv_np).cuda()
    vals = t.from_numpy(v_np).cuda()
    print(f"[server] {n} one-visibility requests on {jobs} job ids", flush=True)
    with migrid.Server() as srv:
        tickets = [srv.submit(j % jobs, uv, vals, gcf, N, S) for j in range(n)]
        srv.wait(tickets[-1])
        for tk in tickets[-100:]:
            srv.wait(tk)
        st = srv.stats()
        assert st["completed"] == n and st["failed"] == 0
        with pytest.raises(migrid.MigridError, match="released"):
            srv.poll(tickets[0])
        with pytest.raises(migrid.MigridError, match="released"):
            srv.wait(tickets[1])
        assert srv.poll(tickets[-1])[0]
        one = ref(uv_np, v_np, g_np, N, S)
        for j in (0, jobs - 1):
            got = srv.read_grid(j)
            count = len(range(j, n, jobs))
            assert np.abs(got - count * one).max() <= 1e-5 * count * np.abs(one).max()


@pytest.mark.parametrize("S", [8, 32])
def test_kept_sort_is_reused_for_repeated_coordinates(S):
    """A job keeps its request's bin sort; later requests with the same uv and
    new values reuse it (only their values are packed) and accumulate as if
    sorted afresh -- also when a batch mixes jobs sorted afresh with jobs
    taking kept sorts.  A request claiming a kept sort its job does not hold
    fails alone."""
    t = torch()
    N, n = 512, 20000
    uv, vals, gcf = inputs(1, N, S, n, seed=31)
    rng = np.random.default_rng(S)
    v2 = rng.normal(size=vals.shape)
    uv_b, vals_b, _ = inputs(2, N, S, 7000, seed=32)
    d = [t.from_numpy(np.ascontiguousarray(x)).cuda() for x in (uv, vals, gcf, v2, 3.0 * v2, uv_b, vals_b)]
    # a long request first keeps the one slot busy, so the next ones share a batch
    uv_l, vals_l, gcf_l = inputs(3, 2048, 64, 2_000_000, seed=33)
    dl = [t.from_numpy(np.ascontiguousarray(x)).cuda() for x in (uv_l, vals_l, gcf_l)]
    t.cuda.synchronize()
    with migrid.Server(arith=migrid.ARITH_NATIVE, max_inflight=1) as srv:
        srv.wait(srv.submit(4, d[0], d[1], d[2], N, S, keep_sort=True))
        tl = srv.submit(9, *dl, 2048, 64)
        ts = [srv.submit(4, d[0], d[3], d[2], N, S, same_uv=True),
              srv.submit(6, d[5], d[6], d[2], N, S)]
        srv.wait(tl)
        for tk in ts:
            srv.wait(tk)
        srv.wait(srv.submit(4, d[0], d[4], d[2], N, S, same_uv=True))
        st = srv.stats()
        assert st["sorts_reused"] == 2 and st["failed"] == 0
        want = ref(uv, vals, gcf, N, S) + 4.0 * ref(uv, v2, gcf, N, S)
        assert O.rel_errors(srv.read_grid(4), want)[0] < 1e-12
        assert O.rel_errors(srv.read_grid(6), ref(uv_b, vals_b, gcf, N, S))[0] < 1e-12
        # no kept sort for job 7, and job 4's is of 20000 visibilities, not 100
        with pytest.raises(migrid.MigridError):
            srv.wait(srv.submit(7, d[0], d[1], d[2], N, S, same_uv=True))
        with pytest.raises(migrid.MigridError):
            srv.wait(srv.submit(4, d[0][:100], d[1][:100], d[2], N, S, same_uv=True))
        srv.wait(srv.submit(4, d[0], d[1], d[2], N, S, same_uv=True))
        want = want + ref(uv, vals, gcf, N, S)
        assert O.rel_errors(srv.read_grid(4), want)[0] < 1e-12
        assert srv.stats()["failed"] == 2
