"""GPU tests of device-resident problems, batched solves and the request server.

Bar: FP64 fields against the oracle to 1e-8 (tight solves); refined served
fields within 1e-5 of the tightly converged oracle; a request solved in a
heterogeneous batch or in the server mix gives the same field as the same
request solved alone (bitwise: per-request tiling and fixed-order
reductions make the arithmetic independent of the mix; the assertion allows
1e-14 relative).
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _request(native, shape, seed, lam=542.0, method="LDR", images=0, theta=0.0, phi=0.0, psi=0.0):
    rng = np.random.default_rng(seed)
    k = (math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))
    e1 = (math.cos(theta) * math.cos(phi), math.cos(theta) * math.sin(phi), -math.sin(theta))
    e2 = (-math.sin(phi), math.cos(phi), 0.0)
    e = tuple(math.cos(psi) * a + math.sin(psi) * b for a, b in zip(e1, e2))
    return native.Request(shape=shape, d=15.0, lam=lam, rho=rng.uniform(0, 1, int(np.prod(shape))), n_k=k,
                          n_e0=e, method=method, images=images, period=(shape[0], shape[1]))


def _oracle_field(O, req, tables, tol=1e-12):
    box = O.OracleFullBox(req.shape, req.d, req.lam, req.rho, n_k=req.n_k, n_e0=req.n_e0, e0=req.e0,
                          material=req.material, nback=req.nback, method=req.method, images=req.images,
                          period=req.period, tables=tables, workers=4)
    box.solve(tol=tol)
    return box.field()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def native():
    from dda_model import native
    return native


def test_problem_field_matches_oracle(native, tables):
    import dda_oracle as O
    _log("[gpu] problem vs oracle")
    req = _request(native, (12, 10, 6), seed=3, theta=0.3, phi=1.0, psi=0.4)
    p = native.Problem(req, precision="fp64", tables=tables)
    info = native.solve([p], tol=1e-12)[0]
    assert info["status"] == "converged"
    assert _rel(p.field(), _oracle_field(O, req, tables)) < 1e-8


def test_problem_fcd_periodic_matches_oracle(native, tables):
    import dda_oracle as O
    _log("[gpu] periodic FCD problem vs oracle")
    req = _request(native, (8, 8, 4), seed=4, lam=600.0, method="FCD", images=4)
    p = native.Problem(req, precision="fp64", tables=tables)
    native.solve([p], tol=1e-12)
    assert _rel(p.field(), _oracle_field(O, req, tables)) < 1e-8


@pytest.mark.parametrize("precond", ["none", "jacobi"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_batch_equals_alone(native, tables, precision, precond):
    _log(f"[gpu] batch vs alone {precision} {precond}")
    reqs = [_request(native, s, seed=10 + i, lam=450 + 40 * i, theta=0.2 * i)
            for i, s in enumerate([(16, 16, 16), (24, 24, 24), (12, 20, 8)])]
    alone = []
    for r in reqs:
        p = native.Problem(r, precision=precision, tables=tables, precond=precond)
        native.solve([p])
        alone.append(p.field())
    ps = [native.Problem(r, precision=precision, tables=tables, precond=precond) for r in reqs]
    infos = native.solve(ps)
    for p, a, inf in zip(ps, alone, infos):
        assert inf["status"] == "converged"
        assert _rel(p.field(), a) <= 1e-14


def test_batched_green_setup_equals_direct(native, tables):
    """Operators built together (one launch per set-up step for all of them,
    as the server's admission rounds do) give the same spectra as operators
    built one by one: the same matvecs, bitwise, with c64 and FP64
    transforms.  The round holds two box shapes, a repeated operator (served
    from the round itself) and an FCD box with images (lattice sum)."""
    _log("[gpu] batched Green's set-up vs direct")
    reqs = [_request(native, (16, 16, 16), seed=1, lam=431.0),
            _request(native, (20, 12, 10), seed=2, lam=512.0, theta=0.2),
            _request(native, (16, 16, 16), seed=3, lam=431.0),
            _request(native, (12, 10, 6), seed=4, lam=601.0, method="FCD", images=2),
            _request(native, (24, 24, 24), seed=5, lam=702.0, method="FCD")]
    rng = np.random.default_rng(9)
    xs = [rng.standard_normal(3 * int(np.prod(r.shape))) + 1j * rng.standard_normal(3 * int(np.prod(r.shape)))
          for r in reqs]
    native.trim_caches()
    batched = native.Problem.create_many(reqs, precision="fp32_refined", tables=tables)
    cached = [p.grid()[2] for p in batched]
    assert cached == [False, False, True, False, False]
    yb = {c64: native.apply(batched, xs, c64=c64) for c64 in (True, False)}
    del batched
    native.trim_caches()
    direct = [native.Problem(r, precision="fp32_refined", tables=tables) for r in reqs]
    assert [p.grid()[2] for p in direct] == [False, False, True, False, False]
    for c64 in (True, False):
        yd = native.apply(direct, xs, c64=c64)
        for a, b in zip(yb[c64], yd):
            assert np.array_equal(a, b)


def test_batched_create_failure_releases_operators(native, tables):
    """A batched creation that fails after queuing operators (an unknown
    method on its third request) raises, builds none of them and leaves
    their keys free: the same operators are then built and give the matvec
    of a one-by-one build."""
    _log("[gpu] batched creation failure")
    good = [_request(native, (12, 10, 6), seed=21, lam=433.0), _request(native, (16, 16, 16), seed=22, lam=517.0)]
    bad = _request(native, (8, 8, 8), seed=23, lam=611.0)
    native.trim_caches()
    descs = [r.desc() for r in good] + [bad.desc()]
    descs[2].method = 7
    arr = (native.ProblemDesc * 3)(*descs)
    hs = (native.ctypes.c_void_p * 3)()
    with pytest.raises(ValueError):
        native.check(native.lib().tdda_problems_create(arr, 3, native.PREC["fp64"], None, None, 0, 0.1, hs))
    again = native.Problem.create_many(good, precision="fp64", tables=tables)
    assert [p.grid()[2] for p in again] == [False, False]
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal(3 * p.n) + 1j * rng.standard_normal(3 * p.n) for p in again]
    ya = native.apply(again, xs)
    del again
    native.trim_caches()
    one = [native.Problem(r, precision="fp64", tables=tables) for r in good]
    for a, b in zip(ya, native.apply(one, xs)):
        assert np.array_equal(a, b)


def test_server_mix_equals_alone_and_oracle(native, tables):
    import dda_oracle as O
    _log("[gpu] server mix vs alone")
    shapes = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (16, 16, 16), (20, 12, 10), (24, 24, 24)]
    reqs = [_request(native, s, seed=100 + i, lam=400 + 60 * i, theta=0.1 * i, phi=0.5 * i, psi=0.3 * i)
            for i, s in enumerate(shapes)]
    srv = native.Server(precision="fp32_refined", max_slots=8, chunk=3)
    for i, r in enumerate(reqs):
        srv.submit(i, r)
    got = {}
    for _ in range(6000):
        for res in srv.poll():
            got[res["id"]] = res
        if len(got) == len(reqs):
            break
        import time
        time.sleep(0.005)
    assert len(got) == len(reqs)
    mix = {i: srv.field(i) for i in got}
    st = srv.stats()
    srv.close()
    _log(f"[gpu] server stats {st}")
    assert st["max_active"] >= 2
    for i, r in enumerate(reqs[:3]):
        s1 = native.Server(precision="fp32_refined", max_slots=1, chunk=5)
        s1.submit(0, r)
        res = []
        while not res:
            res = s1.poll()
        alone = s1.field(0)
        s1.close()
        assert got[i]["status"] == "converged" and res[0]["iterations"] == got[i]["iterations"]
        assert _rel(mix[i], alone) <= 1e-14
    # accuracy of served results against the tightly converged oracle: the
    # refined server stops when its field-error estimate meets tol (1e-5)
    for i in (0, 4):
        e_ref = _oracle_field(O, reqs[i], tables)
        err = _rel(mix[i], e_ref)
        _log(f"[gpu] served request {i}: E err {err:.3e}, estimate {got[i]['err_est']:.3e}, "
             f"true residual {got[i]['error']:.3e}")
        assert got[i]["error"] <= 1e-5
        assert err <= 1e-5
        assert err / 3.0 <= got[i]["err_est"] <= 3.0 * err


def test_server_admission_by_memory(native, tables):
    """A budget below two requests' footprint keeps one request in flight; all finish."""
    _log("[gpu] server memory admission")
    reqs = [_request(native, (24, 24, 24), seed=200 + i) for i in range(3)]
    native.check(native.lib().tdda_trim_caches())
    base = native.lib().tdda_device_bytes()
    # one 24^3 request needs about 13 MB (estimate) and holds about 9 MB: a
    # budget of base + 20 MB fits one request in flight but not two
    srv = native.Server(precision="fp32_refined", max_slots=8, hbm_budget_gb=(base + 20e6) / 2 ** 30, chunk=4)
    for i, r in enumerate(reqs):
        srv.submit(i, r)
    import time
    done = []
    t_end = time.time() + 60
    while len(done) < 3 and time.time() < t_end:
        done += srv.poll()
        time.sleep(0.005)
    st = srv.stats()
    srv.close()
    assert len(done) == 3 and all(d["status"] == "converged" for d in done)
    assert st["max_active"] == 1


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_jacobi_preconditioned_solves(native, tables, precision):
    """Right Jacobi preconditioning (BiCGSTAB on A diag(al)^-1, P = Y / al):
    the residual is the original system's, so a solve to 1e-12 gives the
    oracle's field, and a refined solve to 1e-5 meets the field bound, in
    far fewer iterations than the plain recurrence."""
    import dda_oracle as O
    _log(f"[gpu] jacobi {precision}")
    for i, (shape, method, images) in enumerate([((16, 16, 16), "LDR", 0), ((8, 8, 4), "FCD", 4)]):
        req = _request(native, shape, seed=300 + i, lam=520.0 + 40 * i, method=method, images=images,
                       theta=0.25, phi=0.7, psi=1.3)
        e_ref = _oracle_field(O, req, tables)
        if precision == "fp64":
            p = native.Problem(req, precision="fp64", tables=tables, precond="jacobi")
            info = native.solve([p], tol=1e-12)[0]
            assert info["status"] == "converged"
            assert _rel(p.field(), e_ref) < 1e-9
        its = {}
        for pc in ("none", "jacobi"):
            p = native.Problem(req, precision=precision, tables=tables, precond=pc)
            info = native.solve([p], tol=1e-5)[0]
            its[pc] = info["iterations"]
            err = _rel(p.field(), e_ref)
            _log(f"[gpu] {shape} {method} {precision} {pc}: {info}; E err {err:.3e}")
            assert info["status"] == "converged"
            if precision == "fp32_refined":
                assert err <= 1e-5
            else:
                assert err <= 5e-5
        assert its["jacobi"] < its["none"]


def test_server_jacobi_mix_equals_alone_and_oracle(native, tables):
    import dda_oracle as O
    _log("[gpu] jacobi server mix vs alone")
    shapes = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (20, 12, 10)]
    reqs = [_request(native, s, seed=400 + i, lam=430 + 90 * i, theta=0.1 * i, phi=0.4 * i, psi=0.9 * i)
            for i, s in enumerate(shapes)]
    srv = native.Server(precision="fp32_refined", max_slots=8, chunk=2, precond="jacobi")
    for i, r in enumerate(reqs):
        srv.submit(i, r)
    import time
    got = {}
    t_end = time.time() + 60
    while len(got) < len(reqs) and time.time() < t_end:
        for res in srv.poll():
            got[res["id"]] = res
        time.sleep(0.002)
    assert len(got) == len(reqs)
    mix = {i: srv.field(i) for i in got}
    srv.close()
    for i, r in enumerate(reqs):
        s1 = native.Server(precision="fp32_refined", max_slots=1, chunk=3, precond="jacobi")
        s1.submit(0, r)
        res = []
        while not res:
            res = s1.poll()
            time.sleep(0.001)
        alone = s1.field(0)
        s1.close()
        assert got[i]["status"] == "converged" and res[0]["iterations"] == got[i]["iterations"]
        assert _rel(mix[i], alone) <= 1e-14
        if i in (0, 3):
            err = _rel(mix[i], _oracle_field(O, r, tables))
            _log(f"[gpu] jacobi served {i}: {got[i]['iterations']} its, E err {err:.3e}, estimate {got[i]['err_est']:.3e}")
            assert err <= 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_graph_replay_equals_eager(native, tables, precision):
    """Batched solves replay one recorded BiCGSTAB iteration per step (a HIP
    graph, launched hundreds of times with the host's chunk checks in
    between); the same solves with the launches issued one by one give the
    same iterations and bitwise the same fields, every replay having fed the
    next.  The server's chunk graphs (recorded once a batch is reused, as in
    the tail of a stream) are checked the same way."""
    _log(f"[gpu] graph vs eager {precision}")
    reqs = [_request(native, s, seed=500 + i, lam=470 + 55 * i, theta=0.15 * i, psi=0.4 * i)
            for i, s in enumerate([(16, 16, 16), (20, 12, 10), (24, 24, 24)])]
    out = {}
    for on in (True, False):
        native.set_graph_capture(on)
        try:
            for pc in ("none", "jacobi"):
                ps = [native.Problem(r, precision=precision, tables=tables, precond=pc) for r in reqs]
                infos = native.solve(ps)
                out[(on, pc)] = ([i["iterations"] for i in infos], [p.field() for p in ps])
            srv = native.Server(precision=precision, max_slots=4, chunk=2, precond="jacobi")
            for i, r in enumerate(reqs):
                srv.submit(i, r)
            got = {}
            import time
            t_end = time.time() + 60
            while len(got) < len(reqs) and time.time() < t_end:
                for res in srv.poll():
                    got[res["id"]] = res
                time.sleep(0.002)
            out[(on, "server")] = ([got[i]["iterations"] for i in range(len(reqs))],
                                   [srv.field(i) for i in range(len(reqs))])
            srv.close()
        finally:
            native.set_graph_capture(True)
    for key in ("none", "jacobi", "server"):
        its_g, f_g = out[(True, key)]
        its_e, f_e = out[(False, key)]
        assert its_g == its_e, key
        for a, b in zip(f_g, f_e):
            assert np.array_equal(a, b), key
