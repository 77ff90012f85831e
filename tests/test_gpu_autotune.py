"""Factorisation paths and the autotuner that picks among them.

The batched evaluation can factorise the swarm with one persistent launch
(path 0) or with lock-step launches whose trailing-update GEMM stages 32-deep
(path 1) or 16-deep chunks (paths 2-4, with 64x128, 32x128 or 32x64 update
tiles; path 5 mixes them panel by panel).  With 16-deep chunks the lock-step
schedule is chained by default (two launches per panel, the diagonal chain
handed over between work-groups inside the update launch; FB_CHAIN=0 gives
three launches).  All of them run the same tile routines in the same
per-element accumulation order, so their results must be bit-identical; the
autotuner (first call per swarm size) may then pick any of them.  Each path
is also held to the oracle.
"""
import numpy as np
import pytest

from ferebus_amd import SwarmLikelihood, compile_kernel, fit_mean, synthetic
from test_gpu_parity import RTOL, _check, _lapack, _oracle

pytestmark = pytest.mark.gpu


def _gp(x, y, prog, mfit, S, noise=1e-6):
    return SwarmLikelihood(x, y, prog, mean=mfit, noise=noise, max_swarm=S)


def _problem(n, d, seed):
    rng = np.random.default_rng(seed)
    x = synthetic.alf_features(rng, n, d)
    y = synthetic.bump_target(rng, x, nbumps=5)
    prog = compile_kernel("k1+k2", synthetic.rbf_periodic_kernels(d), d)
    return rng, x, y, prog, fit_mean("constant", x, y)


@pytest.mark.parametrize("n,S", [(300, 5), (641, 9), (1000, 1)])
def test_paths_bit_identical(gpu, monkeypatch, n, S):
    rng, x, y, prog, mfit = _problem(n, 6, seed=n)
    out = {}
    for path in range(5):
        monkeypatch.setenv("FB_FACTOR_PATH", str(path))
        gp = _gp(x, y, prog, mfit, S)
        if path == 0:
            th = rng.uniform(0.0, 1.0, (S, gp.nparams))
        ll, info = gp.swarm_loglik(th)
        ll2, info2 = gp.swarm_loglik(th)          # graph replay
        assert np.array_equal(ll, ll2) and np.array_equal(info, info2)
        assert gp.factor_path(S) == SwarmLikelihood.FACTOR_PATHS[path]
        # the persistent launch was served as scheduled: no lock-step re-run
        assert gp.profile_read("dispatch_fallbacks") == (0.0, 0)
        out[path] = (ll, info)
    print({p: o[1].tolist() for p, o in out.items()})
    for path in range(1, 5):
        assert np.array_equal(out[path][1], out[0][1]), f"path {path}: INFO differs"
        ok = out[0][1] == 0
        assert np.array_equal(out[path][0][ok], out[0][0][ok]), f"path {path}: ll differs"
    oll, oinfo = _oracle(x, y, mfit.value, prog, th, 1e-6, False, "marginal")
    _check(out[0][0], out[0][1], oll, oinfo, RTOL, f"n={n} S={S}",
           _lapack(x, y, mfit.value, prog, th, 1e-6, False, "marginal"))


def test_autotune_picks_and_keeps(gpu, monkeypatch):
    monkeypatch.delenv("FB_FACTOR_PATH", raising=False)
    monkeypatch.delenv("FB_PERSISTENT", raising=False)
    monkeypatch.delenv("FB_AUTOTUNE", raising=False)
    rng, x, y, prog, mfit = _problem(900, 6, seed=3)
    gp = _gp(x, y, prog, mfit, 12)
    th = rng.uniform(0.0, 1.0, (12, gp.nparams))
    assert gp.factor_path(12) == "untuned"
    ll, info = gp.swarm_loglik(th)
    chosen = gp.factor_path(12)
    print("autotuned path for S=12:", chosen)
    assert chosen in SwarmLikelihood.FACTOR_PATHS.values() and chosen != "untuned"
    ll2, info2 = gp.swarm_loglik(th)
    assert gp.factor_path(12) == chosen
    assert np.array_equal(ll, ll2) and np.array_equal(info, info2)
    # a smaller swarm is tuned on its own first call, the gradient call included
    assert gp.factor_path(4) == "untuned"
    l4, g4, i4 = gp.swarm_grad(th[:4])
    assert gp.factor_path(4) != "untuned"
    assert np.array_equal(i4, info[:4]) and np.array_equal(l4[i4 == 0], ll[:4][i4 == 0])
    # every path gives the same bits as the one chosen
    monkeypatch.setenv("FB_FACTOR_PATH", "2")
    gp2 = _gp(x, y, prog, mfit, 12)
    ll3, info3 = gp2.swarm_loglik(th)
    assert np.array_equal(info3, info) and np.array_equal(ll3[info == 0], ll[info == 0])


def test_tuning_cache_file(gpu, monkeypatch, tmp_path):
    """FB_TUNE_CACHE: the first handle's choice is written as a JSON record
    and a new handle of the same order starts with it."""
    import json
    monkeypatch.delenv("FB_FACTOR_PATH", raising=False)
    monkeypatch.delenv("FB_PERSISTENT", raising=False)
    monkeypatch.delenv("FB_AUTOTUNE", raising=False)
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("FB_TUNE_CACHE", str(cache))
    rng, x, y, prog, mfit = _problem(500, 6, seed=8)
    gp = _gp(x, y, prog, mfit, 6)
    th = rng.uniform(0.0, 1.0, (6, gp.nparams))
    ll, info = gp.swarm_loglik(th)
    chosen = gp.factor_path(6)
    recs = json.loads(cache.read_text())
    print(recs)
    assert any(r["swarm"] == 6 and SwarmLikelihood.FACTOR_PATHS[r["path"]] == chosen for r in recs)
    gp2 = _gp(x, y, prog, mfit, 6)
    assert gp2.factor_path(6) == chosen          # known before any evaluation
    ll2, info2 = gp2.swarm_loglik(th)
    assert np.array_equal(ll, ll2) and np.array_equal(info, info2)


def _dup_case(n, dup, src, seed):
    """rbf on 6 features with the nugget in theta; point `dup` repeats point
    `src`, so a negative nugget makes column dup + 1 fail (INFO = dup + 1)
    and a positive one leaves K positive definite."""
    from ferebus_amd import KernelConfig
    rng = np.random.default_rng(seed)
    x = synthetic.alf_features(rng, n, 6)
    x[dup] = x[src]
    y = synthetic.bump_target(rng, x, nbumps=4)
    prog = compile_kernel("k1", {"k1": KernelConfig("k1", "rbf")}, 6)
    return x, y, prog, fit_mean("constant", x, y)


@pytest.mark.parametrize("n,dup,src,S,code", [
    (700, 200, 5, 9, 0),     # fails in the second 64-block of panel 1 (columns 192..255)
    (700, 300, 17, 9, 0),    # fails in the first 64-block of panel 2 (columns 256..319)
    (383, 130, 2, 1, 0),     # single particle, bordered rows in the last panel
    (700, 200, 5, 9, 4),     # the same failures with two role work-groups,
    (700, 300, 17, 9, 4),
    (700, 200, 5, 9, 8),     # with the fused row solve (three and two roles)
    (700, 300, 17, 9, 12),
    (700, 200, 5, 9, 16),    # 8-deep chunks, three work-groups per CU
    (700, 300, 17, 9, 20),
    (700, 200, 5, 9, 32),    # the previous panel's rows solved in the update launch
    (700, 300, 17, 9, 36),   # (three and two roles, 16- and 8-deep chunks)
    (700, 200, 5, 9, 48),
    (700, 300, 17, 9, 52),
    (700, 200, 5, 9, 64),    # 8-deep chunks, four work-groups per CU (alone and ahead)
    (700, 300, 17, 9, 100),
])
def test_chained_schedule_bit_identical(gpu, monkeypatch, n, dup, src, S, code):
    """Two launches per panel (the diagonal chain handed over inside the
    update launch; code 4: two role work-groups, 8 / 12: the row solve fused
    into the update launch, 16 / 20: 8-deep chunks at three work-groups per
    CU) against three: same bits, same INFO, the failing
    particles included (every role must publish on every exit path)."""
    x, y, prog, mfit = _dup_case(n, dup, src, seed=n + dup)
    out = {}
    npan = (n + 1 + 127) // 128
    for chain in (0, 1):
        monkeypatch.setenv("FB_CHAIN", str(chain))
        if code:
            monkeypatch.setenv("FB_FACTOR_PATH", "5")
            monkeypatch.setenv("FB_PANEL_TILES", ",".join([str(code if chain else 0)] * npan))
        else:
            monkeypatch.setenv("FB_FACTOR_PATH", "2")
        gp = SwarmLikelihood(x, y, prog, mean=mfit, noise=1e-6, optimise_noise=True, max_swarm=S)
        th = np.random.default_rng(3).uniform(0.0, 1.0, (S, gp.nparams))
        th[:, -1] = np.where(np.arange(S) % 3 == 1, -1e-8, 1e-4)
        ll, info = gp.swarm_loglik(th)
        ll2, info2 = gp.swarm_loglik(th)          # graph replay: the flags start over
        assert np.array_equal(ll, ll2, equal_nan=True) and np.array_equal(info, info2)
        out[chain] = (ll, info, th)
    print("INFO three launches", out[0][1].tolist(), "two launches", out[1][1].tolist())
    assert np.array_equal(out[0][1], out[1][1])
    ok = out[0][1] == 0
    assert np.array_equal(out[0][0][ok], out[1][0][ok])
    expect = np.where(np.arange(S) % 3 == 1, dup + 1, 0)
    assert np.array_equal(out[1][1], expect)
    th = out[1][2]
    oll, oinfo = _oracle(x, y, mfit.value, prog, th, 1e-6, True, "marginal")
    assert np.array_equal(oinfo, expect)
    _check(out[1][0], out[1][1], oll, oinfo, 1e-8, f"chained n={n}")


@pytest.mark.parametrize("path,tiles", [(3, None), (4, None), (5, "2,0,1,2,0,1,2,0"), (5, "6,4,5,2,4,1,6,5"),
                                        (5, "8,12,8,12,0,8,4,12"), (5, "16,20,16,20,0,16,4,20")])
def test_chained_schedule_tile_shapes(gpu, monkeypatch, path, tiles):
    """The chained update launch with the 32x128 and 32x64 row tiles, the
    per-panel mix, the two-role variant (tile code + 4), the fused row
    solve (8, 12: one launch for the panel) and the 8-deep chunks at three
    work-groups per CU (16, 20): the bits of the three-launch schedule."""
    _, x, y, prog, mfit = _problem(900, 6, seed=path)
    out = {}
    for chain in (0, 1):
        monkeypatch.setenv("FB_CHAIN", str(chain))
        monkeypatch.setenv("FB_FACTOR_PATH", str(path))
        if tiles:
            # the three-launch schedule has no role work-groups: same shapes
            monkeypatch.setenv("FB_PANEL_TILES", tiles if chain else ",".join(str(int(t) & 3) for t in tiles.split(",")))
        gp = _gp(x, y, prog, mfit, 7)
        th = np.random.default_rng(4).uniform(0.0, 1.0, (7, gp.nparams))
        out[chain] = gp.swarm_loglik(th)
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("n,tiles,tail", [(900, "32,32,32,32,32,32,32,32", 0),
                                          (900, "36,52,48,32,36,48,52,32", 0),
                                          (900, "8,32,0,32,16,48,2,32", 0),
                                          (1100, "0,32,32,32,0,32,32,32,36", 3),
                                          (1100, "0,32,32,32,32,32,32,32,32", 3),
                                          (1100, "0,32,32,32,32,12,12,12,0", 3),
                                          (1100, "32,32,4,32,48,0,32,52,32", 2),
                                          (250, "32,32,32", 0),
                                          (900, "64,96,100,68,64,96,4,100", 0),
                                          (1100, "64,64,96,64,64,96,96,68,96", 3)])
def test_ahead_row_solve_bits(gpu, monkeypatch, n, tiles, tail):
    """Update launches that also solve the previous panel's rows (tile code
    + 32, alone, with two role work-groups (+ 4) or 8-deep chunks at three
    or four work-groups per CU (+ 16, + 64); on the first panel and after a
    fused one (+ 8) the bit is ignored), mixed with the other variants, after tail pre-updates (panels 5-8 of 9 start
    from the pre-updated columns) and in the launch that would carry them
    (then it carries none), and the autotuner's second uniform candidate
    (ahead up to the tail pre-update, from there the row solve fused into
    each update launch with two role work-groups): the bits of the
    three-launch schedule."""
    _, x, y, prog, mfit = _problem(n, 6, seed=n + tail)
    out = {}
    for chain in (0, 1):
        monkeypatch.setenv("FB_CHAIN", str(chain))
        monkeypatch.setenv("FB_TAIL", str(tail))
        monkeypatch.setenv("FB_FACTOR_PATH", "5")
        monkeypatch.setenv("FB_PANEL_TILES", tiles if chain else ",".join(str(int(t) & 3) for t in tiles.split(",")))
        gp = _gp(x, y, prog, mfit, 9)
        th = np.random.default_rng(n).uniform(0.0, 1.0, (9, gp.nparams))
        out[chain] = gp.swarm_loglik(th)
        # the graph replay of the same evaluation
        out[chain + 2] = gp.swarm_loglik(th)
    print("info", out[1][1].tolist())
    for k in (1, 2, 3):
        assert np.array_equal(out[0][1], out[k][1]) and np.array_equal(out[0][0], out[k][0], equal_nan=True)


@pytest.mark.parametrize("tail,lookahead", [(1, 0), (5, 0), (7, 0), (0, 1), (3, 1), (7, 1)])
def test_tail_pre_update_bits(gpu, monkeypatch, tail, lookahead):
    """The chained schedule with 1, 5 and 7 trailing panels pre-updated by the
    launch before them (9 panels here; 7 leaves panel 1 carrying all of
    them), alone or with every launch also pre-updating the next panel
    (look-ahead): the bits of the three-launch schedule."""
    _, x, y, prog, mfit = _problem(1100, 6, seed=tail + 10 * lookahead)
    out = {}
    for chain in (0, 1):
        monkeypatch.setenv("FB_CHAIN", str(chain))
        monkeypatch.setenv("FB_TAIL", str(tail))
        monkeypatch.setenv("FB_LOOKAHEAD", str(lookahead))
        monkeypatch.setenv("FB_FACTOR_PATH", "2")
        gp = _gp(x, y, prog, mfit, 5)
        th = np.random.default_rng(6).uniform(0.0, 1.0, (5, gp.nparams))
        out[chain] = gp.swarm_loglik(th)
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][0], out[1][0])


def test_stream_groups_tuned_and_bits(gpu, monkeypatch, tmp_path):
    """The autotuner also picks how many stream groups the swarm's
    factorisations run as (1-4, after the path); the tuning cache keeps the
    choice; every group count gives the bits of one group (graph replay
    included); and a swarm above 160 particles split into groups of at most
    160 runs the chained schedule in each, with the bits of single-particle
    evaluations."""
    import json
    for k in ("FB_FACTOR_PATH", "FB_PERSISTENT", "FB_AUTOTUNE", "FB_STREAMS"):
        monkeypatch.delenv(k, raising=False)
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("FB_TUNE_CACHE", str(cache))
    rng, x, y, prog, mfit = _problem(1200, 6, seed=31)
    gp = _gp(x, y, prog, mfit, 24)
    th = rng.uniform(0.0, 1.0, (24, gp.nparams))
    assert gp.factor_groups(24) == 1
    ll, info = gp.swarm_loglik(th)
    G = gp.factor_groups(24)
    print("tuned:", gp.factor_path(24), "groups", G)
    assert 1 <= G <= 4
    recs = json.loads(cache.read_text())
    assert any(r["swarm"] == 24 and r["groups"] == G for r in recs)
    gp2 = _gp(x, y, prog, mfit, 24)
    assert gp2.factor_groups(24) == G and gp2.factor_path(24) == gp.factor_path(24)
    monkeypatch.delenv("FB_TUNE_CACHE")
    for streams in (1, 2, 3, 4):
        monkeypatch.setenv("FB_STREAMS", str(streams))
        monkeypatch.setenv("FB_FACTOR_PATH", "2")
        g = _gp(x, y, prog, mfit, 24)
        for _ in range(2):                     # eager, then the graph replay
            l, i = g.swarm_loglik(th)
            assert g.factor_groups(24) == streams
            assert np.array_equal(i, info) and np.array_equal(l, ll, equal_nan=True), f"{streams} groups"
        g.close()
    monkeypatch.setenv("FB_STREAMS", "2")
    _, x3, y3, prog3, mfit3 = _problem(300, 6, seed=11)
    g = _gp(x3, y3, prog3, mfit3, 170)
    th3 = np.random.default_rng(7).uniform(0.0, 1.0, (170, g.nparams))
    l3, i3 = g.swarm_loglik(th3)
    pl, pinfo = g.particle_loglik(th3[::17])
    assert np.array_equal(pinfo, i3[::17]) and np.array_equal(pl[pinfo == 0], l3[::17][pinfo == 0])


@pytest.mark.parametrize("S", [160, 170])
def test_swarm_beyond_chain_limit(gpu, monkeypatch, S):
    """Swarms above 160 particles take the three-launch schedule (the
    chained one needs its waiting work-groups to leave room for the
    producers); 160 particles (the most waiting work-groups the chained
    schedule allows) and 170 give the bits of single-particle evaluations."""
    monkeypatch.setenv("FB_FACTOR_PATH", "2")
    _, x, y, prog, mfit = _problem(300, 6, seed=11)
    gp = _gp(x, y, prog, mfit, S)
    th = np.random.default_rng(7).uniform(0.0, 1.0, (S, gp.nparams))
    ll, info = gp.swarm_loglik(th)
    pl, pinfo = gp.particle_loglik(th[::17])
    assert np.array_equal(pinfo, info[::17]) and np.array_equal(pl[pinfo == 0], ll[::17][pinfo == 0])
    assert (info == 0).sum() > S // 2


def test_chain_wait_bound_reported(gpu, monkeypatch):
    """A hand-over wait of the chained schedule that reaches its bound ends
    the kernel normally and the evaluation returns an error instead of a
    result: forced from the host with a bound of 0 ticks (the consumer gives
    up at its first poll, before its producer has factored its block).  A
    handle with the default bound then evaluates normally."""
    monkeypatch.setenv("FB_FACTOR_PATH", "2")
    _, x, y, prog, mfit = _problem(400, 6, seed=12)
    monkeypatch.setenv("FB_CHAIN_WAIT_TICKS", "0")
    gp = _gp(x, y, prog, mfit, 4)
    th = np.random.default_rng(8).uniform(0.0, 1.0, (4, gp.nparams))
    with pytest.raises(RuntimeError, match="bound"):
        gp.swarm_loglik(th)
    gp.close()
    monkeypatch.delenv("FB_CHAIN_WAIT_TICKS")
    gp = _gp(x, y, prog, mfit, 4)
    ll, info = gp.swarm_loglik(th)
    assert (info == 0).all() and np.isfinite(ll).all()


def test_chained_schedule_stream_groups(gpu, monkeypatch):
    """The swarm split into groups on concurrent streams (FB_STREAMS), each
    group running its own chained schedule and hand-overs: the bits of one
    group, graph replay included."""
    _, x, y, prog, mfit = _problem(700, 6, seed=21)
    monkeypatch.setenv("FB_FACTOR_PATH", "2")
    out = {}
    for streams in (1, 3):
        monkeypatch.setenv("FB_STREAMS", str(streams))
        gp = _gp(x, y, prog, mfit, 11)
        th = np.random.default_rng(9).uniform(0.0, 1.0, (11, gp.nparams))
        ll, info = gp.swarm_loglik(th)
        ll2, info2 = gp.swarm_loglik(th)
        assert np.array_equal(ll, ll2) and np.array_equal(info, info2)
        out[streams] = (ll, info)
        gp.close()
    assert np.array_equal(out[1][1], out[3][1]) and np.array_equal(out[1][0], out[3][0])


def test_retries_do_not_tune(gpu):
    """fb_pso_retry evaluates swarms of re-randomised particles whose count
    changes every iteration: none of them may start the autotuner (it times
    ~40 evaluations per new size); they reuse the nearest tuned size's path
    and give the bits of an ordinary evaluation of the same positions."""
    from ferebus_amd.pso import DeviceSwarm
    w = synthetic.water(n=300, seed=2)
    prog = compile_kernel(w.kernel, w.kernels, 3)
    gp = SwarmLikelihood(w.x, w.y, prog, mean="constant", noise=1e-6, max_swarm=16)
    rng = np.random.default_rng(4)
    sw = DeviceSwarm(gp, 16, 0.7, 1.4, 1.4)
    sw.init(rng.uniform(0, 1, (16, 3)), rng.uniform(-1, 1, (16, 3)))
    sw.step(rng.random((16, 3)), rng.random((16, 3)))
    _, tuned = gp.profile_read("tune_evals")
    assert tuned > 0
    evaluated = []
    for k in (1, 2, 3, 5, 7, 11, 16):
        idx = rng.choice(16, k, replace=False)
        xr = rng.uniform(0, 1, (k, 3))
        c, info = sw.retry(idx, xr, rng.uniform(-1, 1, (k, 3)))
        evaluated.append((xr, c, info))
        _, n = gp.profile_read("tune_evals")
        assert n == tuned, f"retry of {k} particles started the autotuner ({n - tuned} timed evaluations)"
    sw.close()
    for xr, c, info in evaluated:
        ll, i2 = gp.swarm_loglik(xr)
        assert np.array_equal(info, i2)
        assert np.array_equal(c[info == 0], -ll[i2 == 0])
