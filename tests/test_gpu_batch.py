"""GPU tests of the batched form (jw_batch_*): trajectories of uneven sizes,
each against the same trajectory evolved alone with the matching Philox key.

With the same intermediate layout the batched kernels do, per trajectory, the
same arithmetic as the single-lattice ones, so the results agree to the last
bit (checked at 1e-13).  The step graph is replayed several times with
downloads and eager work in between, each replay checked against the eager
single-lattice result.  With the tuner's layout the agreement is to rounding
(FP32: 2e-6 after a few steps).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def frob_max(a, b):
    d = (a - b).reshape(-1, 9)
    return np.sqrt((abs(d) ** 2).sum(axis=1)).max()


def mv_states(jw, groups, Ny=4):
    """One MV start per trajectory (own seed), as host arrays, and per-N contexts."""
    ctxs, states = {}, []
    t = 0
    for N, count in groups:
        if N not in ctxs:
            ctxs[N] = jw.Context(N)
        c = ctxs[N]
        for _ in range(count):
            U = c.field()
            c.philox(100 + t, 0, 0)
            c.mv_init(U, Ny, 1.0, 0.25, 1.0)
            states.append((N, U.download()))
            t += 1
    return ctxs, states


def single(jw, c, U0, fp32, seed, traj, step0, nsteps, coupling=3, ds=1e-4, al=None, reunit=10):
    U = c.field(9, U0)
    c.philox(seed, traj, step0)
    if fp32:
        c.evolve_mixed(U, coupling, 1, ds, al, nsteps=nsteps, reunit_every=reunit)
    else:
        c.evolve(U, coupling, 1, ds, al, nsteps=nsteps)
    return U.download()


@pytest.mark.parametrize("fp32", [1, 0])
def test_batch_matches_single_trajectories(jw, fp32):
    groups = [(64, 3), (128, 2)]
    ctxs, states = mv_states(jw, groups)
    b = jw.Batch(groups, fp32=fp32)
    assert len(b) == 5
    for g in range(len(groups)):
        b.set_layout(g, 0)
    for t, (N, U0) in enumerate(states):
        b.set(t, ctxs[N].field(9, U0))
    b.philox(7, 5, 0)
    # 25 steps: the 10-step graph twice, then 5 steps launched directly
    b.evolve(3, 1, 1e-4, None, nsteps=25, reunit_every=10)
    b.sync()
    for t, (N, U0) in enumerate(states):
        c = ctxs[N]
        got = c.field()
        b.get(t, got)
        ref = single(jw, c, U0, fp32, 7, 5 + t, 0, 25)
        assert frob_max(got.download(), ref) <= 1e-13, t
        assert frob_max(ref, U0) > 1e-3
    b.close()
    for c in ctxs.values():
        c.close()


def test_batch_graph_replays(jw):
    """Three evolutions of 10 steps (one graph replay each) with reads and eager
    single-lattice work in between: every replay matches."""
    import math
    groups = [(64, 2), (256, 1)]
    ctxs, states = mv_states(jw, groups)
    al = jw.alpha(lambda_L=0.02 * 64, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7)))
    b = jw.Batch(groups, fp32=True)
    for g in range(len(groups)):
        b.set_layout(g, 0)
    for t, (N, U0) in enumerate(states):
        b.set(t, ctxs[N].field(9, U0))
    b.philox(3, 0, 0)
    cur = [U0 for _, U0 in states]
    for rep in range(3):
        b.evolve(4, 1, 5e-5, al, nsteps=10, reunit_every=10)
        for t, (N, _) in enumerate(states):
            c = ctxs[N]
            # the single-lattice reference for this replay, from the previous state
            ref = single(jw, c, cur[t], 1, 3, t, 10 * rep, 10, coupling=4, ds=5e-5, al=al)
            got = c.field()
            b.get(t, got)
            g = got.download()
            assert frob_max(g, ref) <= 1e-13, (rep, t)
            cur[t] = ref
    b.close()
    for c in ctxs.values():
        c.close()


@pytest.mark.parametrize("fp32", [1, 0])
def test_batch_back_to_back_alpha_change(jw, fp32):
    """Two evolutions with different running-coupling parameters issued back to back
    (no sync in between: the second call rebuilds the coupling tables the first
    call's graph reads) give the same trajectories as the same sequence with a sync
    in between."""
    import math
    groups = [(128, 2), (64, 1)]
    ctxs, states = mv_states(jw, groups)
    al1 = jw.alpha(lambda_L=0.02 * 128, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7)))
    al2 = jw.alpha(lambda_L=0.05 * 128, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.5)))
    out = []
    for synced in (False, True):
        b = jw.Batch(groups, fp32=fp32)
        for g in range(len(groups)):
            b.set_layout(g, 0)
        for t, (N, U0) in enumerate(states):
            b.set(t, ctxs[N].field(9, U0))
        b.philox(5, 0, 0)
        b.evolve(4, 1, 5e-5, al1, nsteps=20, reunit_every=10)
        if synced:
            b.sync()
        b.evolve(4, 1, 5e-5, al2, nsteps=20, reunit_every=10)
        b.sync()
        res = []
        for t, (N, _) in enumerate(states):
            got = ctxs[N].field()
            b.get(t, got)
            res.append(got.download())
        b.close()
        out.append(res)
    for t in range(len(states)):
        assert np.array_equal(out[0][t], out[1][t]), t
        assert frob_max(out[0][t], states[t][1]) > 1e-3
    for c in ctxs.values():
        c.close()


def test_batch_tuned_layout(jw):
    groups = [(256, 2), (512, 1), (64, 4)]
    ctxs, states = mv_states(jw, groups, Ny=2)
    b = jw.Batch(groups, fp32=True)
    for t, (N, U0) in enumerate(states):
        b.set(t, ctxs[N].field(9, U0))
    b.philox(11, 0, 0)
    b.evolve(3, 1, 1e-4, None, nsteps=12, reunit_every=10)
    b.sync()
    for g in range(len(groups)):
        assert b.layout(g) in (0, 2, 4, 8, 16)
        assert b.fft_mode(g) in (0, 1) and (b.fft_mode(g) == 0 or b.layout(g) == 0)
    gi = [g for g, (N, count) in enumerate(groups) for _ in range(count)]
    for t, (N, U0) in enumerate(states):
        c = ctxs[N]
        got = c.field()
        b.get(t, got)
        c.set_mixed_fft(b.fft_mode(gi[t]))          # the tuner may pick the matrix-core stages
        ref = single(jw, c, U0, 1, 11, t, 0, 12)
        c.set_mixed_fft(0)
        assert frob_max(got.download(), ref) <= 2e-6, t
    b.close()
    for c in ctxs.values():
        c.close()


def test_batch_rejects(jw):
    with pytest.raises(jw.JwError):
        jw.Batch([(32, 2)])
    with pytest.raises(jw.JwError):
        jw.Batch([(64, 0)])
    b = jw.Batch([(64, 1)])
    with pytest.raises(jw.JwError):
        b.evolve(3, 0, 1e-4, None, 1)          # LINEAR kernel
    with pytest.raises(jw.JwError):
        b.evolve(2, 1, 1e-4, None, 1)          # HATTA
    c = jw.Context(128)
    with pytest.raises(jw.JwError):
        b.set(0, c.field())                    # wrong N
    with pytest.raises(jw.JwError):
        b.set(1, c.field())                    # no such trajectory
    b.close()
    c.close()


@pytest.mark.parametrize("fp32", [1, 0])
@pytest.mark.parametrize("kspace", [0, 1])
def test_batch_one_launch_matches_single(jw, fp32, kspace):
    """One launch per kernel class over every size group (jw_batch_set_launch(1)):
    the same arithmetic per trajectory as a single lattice with row-major planes, so
    every trajectory agrees with it to 1e-13, with several rows or column strips of
    the small groups in one workgroup (64^2, 128^2 beside 256^2 and 2048^2)."""
    import math
    groups = [(64, 3), (128, 2), (256, 1), (2048, 1)]
    print(f"batch: one launch per kernel class, fp32={fp32}, momentum noise={kspace}", flush=True)
    ctxs, states = mv_states(jw, groups, Ny=2)
    b = jw.Batch(groups, fp32=fp32)
    b.set_launch(1)
    assert b.launch() == 1
    if kspace:
        b.set_noise_space(jw.NOISE_MOMENTUM)
    for t, (N, U0) in enumerate(states):
        b.set(t, ctxs[N].field(9, U0))
    b.philox(5, 9, 0)
    coupling, ds = (4, 5e-5) if kspace else (3, 1e-4)
    al = lambda N: jw.alpha(lambda_L=0.02 * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7))) if coupling == 4 \
        else None
    # the batch passes one alpha; each group's coupling table is built for its own N from it
    b.evolve(coupling, 1, ds, al(64), nsteps=12, reunit_every=10)
    b.sync()
    for t, (N, U0) in enumerate(states):
        c = ctxs[N]
        if kspace:
            c.set_noise_space(jw.NOISE_MOMENTUM)
        got = c.field()
        b.get(t, got)
        ref = single(jw, c, U0, fp32, 5, 9 + t, 0, 12, coupling=coupling, ds=ds, al=al(64))
        c.set_noise_space(jw.NOISE_POSITION)
        assert frob_max(got.download(), ref) <= 1e-13, (t, N)
        assert frob_max(ref, U0) > 1e-4
    b.close()
    for c in ctxs.values():
        c.close()


def test_batch_launch_tuning(jw):
    """The launch mode is timed on the first evolution; either choice evolves every
    trajectory as a single lattice does (FP32 rounding apart from the layouts)."""
    groups = [(256, 4), (512, 2)]
    ctxs, states = mv_states(jw, groups, Ny=2)
    b = jw.Batch(groups, fp32=True)
    assert b.launch() == -1
    for t, (N, U0) in enumerate(states):
        b.set(t, ctxs[N].field(9, U0))
    b.philox(3, 0, 0)
    b.evolve(3, 1, 1e-4, None, nsteps=10, reunit_every=10)
    b.sync()
    print(f"batch: tuned launch mode {b.launch()}, layouts {[b.layout(g) for g in range(2)]}", flush=True)
    assert b.launch() in (0, 1)
    gi = [g for g, (N, count) in enumerate(groups) for _ in range(count)]
    for t, (N, U0) in enumerate(states):
        c = ctxs[N]
        got = c.field()
        b.get(t, got)
        c.set_mixed_fft(b.fft_mode(gi[t]))
        ref = single(jw, c, U0, 1, 3, t, 0, 10)
        c.set_mixed_fft(0)
        assert frob_max(got.download(), ref) <= 2e-6, t
    with pytest.raises(jw.JwError):
        b.set_launch(2)
    b.close()
    for c in ctxs.values():
        c.close()
