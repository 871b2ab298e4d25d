"""GPU checks of the learned initial guess: the BF16 MFMA forward pass and
gradients against the numpy oracle (with the same bf16 rounding points), loss
decrease under training, and the gain of the trained network over the
predictors without it on the solver's own laser-scan trajectory."""
import numpy as np
import pytest

from oracle import cnn_oracle as N
from tests.test_cnn_oracle import init, sample

pytestmark = pytest.mark.gpu


def to_dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a).ravel(), dtype=torch.float64, device="cuda")


def make_net(M, shape, degree=1, rho=1, **kw):
    """Network for a (nz, ny, nx) DoF grid of Q_degree."""
    cells = tuple((n - 1) // degree for n in (shape[2], shape[1], shape[0]))
    return M.LearnedPredictor(cells, degree=degree, coarsening=rho, **kw), N.transfers(degree, cells, rho)


@pytest.mark.parametrize("shape,degree,rho", [((5, 6, 7), 1, 1), ((4, 9, 21), 1, 1), ((17, 13, 9), 2, 2),
                                              ((13, 25, 13), 3, 4), ((9, 13, 17), 4, 1), ((9, 17, 17), 4, 2),
                                              ((17, 17, 33), 2, 8)])
def test_forward_matches_oracle(cuda, shape, degree, rho):
    torch = cuda
    import meltpooldg_amd as M
    Tnm1, Tn, q, z, _ = sample(shape)
    net, tr = make_net(M, shape, degree, rho, qscale=0.7)
    w = init().astype(np.float32)
    net.set_weights(w)
    out = torch.empty(Tn.size, dtype=torch.float64, device="cuda")
    net.predict(to_dev(torch, Tnm1), to_dev(torch, Tn), to_dev(torch, q), to_dev(torch, z), out)
    torch.cuda.synchronize()
    ref = N.predict(w, Tnm1, Tn, q, z, qscale=0.7, bf=True, tr=tr).ravel()
    lin = (2 * Tn - Tnm1).ravel()
    # compare the learned correction (prediction minus linear extrapolation)
    got, exp = out.cpu().numpy() - lin, ref - lin
    assert np.linalg.norm(got - exp) <= 2e-2 * np.linalg.norm(exp)


@pytest.mark.parametrize("shape,degree,rho", [((6, 7, 18), 1, 1), ((17, 13, 9), 2, 2), ((9, 17, 17), 4, 2),
                                              ((13, 25, 13), 3, 4)])
def test_gradients_match_oracle(cuda, shape, degree, rho):
    torch = cuda
    import meltpooldg_amd as M
    Tnm1, Tn, q, z, Tnp1 = sample(shape, seed=2)
    net, tr = make_net(M, shape, degree, rho, qscale=0.5)
    w = init(3).astype(np.float32)
    net.set_weights(w)
    loss, g = net.gradients(*[to_dev(torch, a) for a in (Tnm1, Tn, q, z, Tnp1)])
    ref_loss, ref = N.loss_and_gradients(w, Tnm1, Tn, q, z, Tnp1, qscale=0.5, bf=True, tr=tr)
    assert abs(loss - ref_loss) <= 2e-2 * ref_loss
    o = 0
    for ci, co in zip(N.CIN, N.COUT):
        n = co * ci * 27 + co
        a, b = g[o:o + n].astype(np.float64), ref[o:o + n]
        o += n
        assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b)
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.995


def test_create_rejects_what_the_transfers_cannot_carry(cuda):
    import meltpooldg_amd as M
    with pytest.raises(M.MpdgError):
        M.LearnedPredictor((1024, 2, 2), degree=1, coarsening=2)  # 1025 DoFs along x
    with pytest.raises(M.MpdgError):
        M.LearnedPredictor((12, 12, 12), degree=1, coarsening=3)
    M.LearnedPredictor((1024, 2, 2), degree=1, coarsening=1)


def test_base_guess_is_what_the_network_corrects(cuda):
    """With an explicit base guess T_b the prediction is T_b plus the same learned correction;
    the linear extrapolation as T_b reproduces the default, and the train step's loss (taken
    before its update) is that of the target T_{n+1} - T_b."""
    torch = cuda
    import meltpooldg_amd as M
    shape = (9, 13, 17)
    Tnm1, Tn, q, z, Tnp1 = sample(shape, seed=8)
    net, _ = make_net(M, shape, degree=2, rho=2, qscale=0.5, seed=5)
    d = [to_dev(torch, a) for a in (Tnm1, Tn, q, z, Tnp1)]
    out0, out1, out2 = (torch.empty_like(d[1]) for _ in range(3))
    lin = 2 * d[1] - d[0]
    other = lin + 0.3 * torch.sin(torch.arange(lin.numel(), device="cuda", dtype=torch.float64))
    net.predict(*d[:4], out0)
    net.predict(*d[:4], out1, base=lin)
    net.predict(*d[:4], out2, base=other)
    torch.cuda.synchronize()
    assert torch.allclose(out1, out0, rtol=0, atol=1e-12 * float(out0.abs().max()))
    assert torch.allclose(out2 - other, out0 - lin, rtol=0, atol=1e-12 * float(out0.abs().max()))
    w = net.get_weights()
    l_default = net.train_step(*d, 1e-3)
    net.set_weights(w)
    l_lin = net.train_step(*d, 1e-3, base=lin)
    net.set_weights(w)
    # T_{n+1} - other differs from T_{n+1} - lin, so the loss differs
    l_other = net.train_step(*d, 1e-3, base=other)
    assert abs(l_lin - l_default) <= 1e-9 * l_default
    assert abs(l_other - l_default) > 1e-3 * l_default


def test_training_reduces_loss(cuda):
    torch = cuda
    import meltpooldg_amd as M
    shape = (8, 8, 16)
    Tnm1, Tn, q, z, Tnp1 = sample(shape, seed=4)
    net, _ = make_net(M, shape, qscale=0.5, seed=7)
    d = [to_dev(torch, a) for a in (Tnm1, Tn, q, z, Tnp1)]
    first = net.train_step(*d, 1e-3)
    for _ in range(60):
        last = net.train_step(*d, 1e-3)
    assert last < 0.2 * first


def test_training_is_bitwise_repeatable(cuda):
    """Two networks with the same seed trained on the same samples end with identical weights and
    losses: the weight gradient sums its wavefronts' partials in a fixed order (no atomics)."""
    torch = cuda
    import meltpooldg_amd as M
    shape = (17, 17, 33)
    runs = []
    for _ in range(2):
        net, _ = make_net(M, shape, degree=2, rho=2, qscale=0.5, seed=7)
        losses = []
        for s in range(3):
            d = [to_dev(torch, a) for a in sample(shape, seed=10 + s)]
            for _ in range(4):
                losses.append(net.train_step(*d, 2e-3))
        runs.append((losses, net.get_weights()))
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])


def test_output_scale_is_set_by_the_first_sample(cuda):
    """The first train step sets the output scale to the target's rms in units of rms(z), and
    predictions use it as the oracle does."""
    torch = cuda
    import meltpooldg_amd as M
    shape = (9, 13, 17)
    Tnm1, Tn, q, z, Tnp1 = sample(shape, seed=6)
    net, tr = make_net(M, shape, degree=2, rho=2, qscale=0.5, seed=3)
    net.train_step(*[to_dev(torch, a) for a in (Tnm1, Tn, q, z, Tnp1)], 1e-3)
    _, sz = N.features(*[N._net(a, tr) for a in (Tnm1, Tn, q, z)], 0.5)
    a, t, tp = (N._net(x, tr) for x in (Tnm1, Tn, Tnp1))
    expect = np.sqrt(np.mean(((tp - (2 * t - a)) / sz) ** 2))
    assert abs(net.output_scale - expect) <= 1e-9 * expect
    w = net.get_weights()
    out = torch.empty(Tn.size, dtype=torch.float64, device="cuda")
    net.predict(to_dev(torch, Tnm1), to_dev(torch, Tn), to_dev(torch, q), to_dev(torch, z), out)
    torch.cuda.synchronize()
    ref = N.predict(w, Tnm1, Tn, q, z, qscale=0.5, bf=True, tr=tr, oscale=net.output_scale).ravel()
    lin = (2 * Tn - Tnm1).ravel()
    got, exp = out.cpu().numpy() - lin, ref - lin
    assert np.linalg.norm(got - exp) <= 2e-2 * np.linalg.norm(exp)


def test_trained_network_beats_the_predictors_without_it(cuda):
    """Q2 on 16^3 cells, laser scan: train on the pairs of steps 2..40, evaluate steps 41..60
    from the same trajectory state. The learned guess (correcting the linear extrapolation)
    starts CG closer to the solution than linear extrapolation (its base, and the untrained
    network's guess) and than the previous step, and needs no more iterations."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan, generate_trajectory, train
    k, cells, dt = 2, 16, 1e-3
    h = 1.0 / cells
    _, Ts, qs, zs, _ = generate_trajectory(k, cells, 60, dt=dt, precond="gmg32", predictor=M.PREDICTOR_LINEAR)
    net = M.LearnedPredictor((cells,) * 3, degree=k, qscale=dt / h, seed=11)
    losses = train(net, Ts[:41], qs[:40], zs[:40], epochs=20, lr=2e-3)
    res = {}
    for name, pred, sm in (("none", M.PREDICTOR_NONE, 0), ("linear", M.PREDICTOR_LINEAR, 0),
                           ("learned", M.PREDICTOR_LEARNED, 0), ("learned_smoothed", M.PREDICTOR_LEARNED, 2)):
        case = LaserScan(k, cells, dt=dt, predictor=pred, learned=net if pred == M.PREDICTOR_LEARNED else None,
                         guess_smoothing=sm, extrapolation_order=1)
        # restart from the trajectory at step 39; one replayed step rebuilds the history (T_39, T_40)
        r0s, its, fbs = [], [], []
        case.t = 39 * dt
        case.heat.set_initial_condition(Ts[39])
        case.step()
        for _ in range(20):
            it, _, (fb, r0, r) = case.step()
            r0s.append(r0)
            its.append(it)
            fbs.append(fb)
        res[name] = (float(np.mean(np.log10(r0s))), float(np.mean(its)), int(sum(fbs)))
    print("training loss", losses[0], losses[-1], "log10 r0 / iterations / fallbacks", res)
    assert losses[-1] < 0.5 * losses[0]
    assert res["learned"][2] == 0
    assert res["learned"][0] < res["linear"][0] - 0.3 and res["linear"][0] < res["none"][0]
    assert res["learned"][1] <= res["linear"][1]
    # two Chebyshev steps of the finest multigrid level on an accepted guess lower its residual
    assert res["learned_smoothed"][2] == 0 and res["learned_smoothed"][0] < res["learned"][0]


@pytest.mark.parametrize("order", [1, 2])
def test_learned_guess_beats_its_base_predictor_q2_32(cuda, order):
    """Q2 on 32^3 cells (274,625 DoFs), laser scan with FP32 multigrid: the network is trained
    on the solver's own steps 2..80 to correct the extrapolation of the given order (its base
    guess), then steps 81..100 are replayed from the same state with that extrapolation alone
    and with the network on top. The learned guess starts at a smaller residual than its base
    on average and in every step (the guess check never falls back), and needs fewer
    iterations on average."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan, generate_trajectory, train
    k, cells, dt = 2, 32, 1e-3
    pred = M.PREDICTOR_LINEAR if order == 1 else M.PREDICTOR_EXTRAPOLATION
    _, Ts, qs, zs, _, bs = generate_trajectory(k, cells, 100, dt=dt, precond="gmg32", predictor=pred,
                                               extrapolation_order=order, pre_smoother_degree=1, with_base=True)
    net = M.LearnedPredictor((cells,) * 3, degree=k, qscale=dt * cells, seed=11)
    losses = train(net, Ts[:81], qs[:80], zs[:80], epochs=12, lr=2e-3, bs=bs[:80])
    res = {}
    for name, p in (("base", pred), ("learned", M.PREDICTOR_LEARNED)):
        case = LaserScan(k, cells, dt=dt, predictor=p, learned=net if p == M.PREDICTOR_LEARNED else None,
                         extrapolation_order=order, pre_smoother_degree=1)
        # the replayed steps before 81 rebuild the history the extrapolation uses
        case.t = (80 - order) * dt
        case.heat.set_initial_condition(Ts[80 - order])
        for _ in range(order):
            case.step()
        r0s, its, fbs = [], [], []
        for _ in range(20):
            it, _, (fb, r0, r) = case.step()
            assert r <= 1.0001e-10
            r0s.append(r0)
            its.append(it)
            fbs.append(fb)
        res[name] = (np.log10(r0s), np.array(its), int(sum(fbs)))
    lb, ib, _ = res["base"]
    ll, il, fl = res["learned"]
    print("order", order, "loss", losses[0], losses[-1], "log10 r0 base/learned %.2f / %.2f" % (lb.mean(), ll.mean()),
          "iterations base/learned %.2f / %.2f" % (ib.mean(), il.mean()), "fallbacks", fl)
    assert losses[-1] < 0.5 * losses[0]
    assert fl == 0
    assert ll.mean() < lb.mean() - 0.3 and np.all(ll < lb)
    assert il.mean() < ib.mean()


def test_rejected_network_is_paused_and_leaves_the_base_run_unchanged(cuda):
    """A network whose guesses never pass the guess check (all weights zero: the guess is its
    base, whose residual is not smaller) is paused for guess_pause_steps steps after
    guess_pause_after consecutive rejections, then tried again. Every step then uses the base
    guess, so the run is bitwise that of the linear predictor alone."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan
    k, cells, dt, n_steps = 2, 8, 1e-3, 12
    net = M.LearnedPredictor((cells,) * 3, degree=k, qscale=dt * cells, seed=3)
    net.set_weights(np.zeros_like(net.get_weights()))
    runs = {}
    for name, pred in (("linear", M.PREDICTOR_LINEAR), ("learned", M.PREDICTOR_LEARNED)):
        case = LaserScan(k, cells, dt=dt, predictor=pred, learned=net if pred == M.PREDICTOR_LEARNED else None,
                         extrapolation_order=1, guess_pause_after=2, guess_pause_steps=3)
        sols, status = [], []
        for _ in range(n_steps):
            _, _, (st, r0, r) = case.step()
            assert r <= 1.0001e-10
            sols.append(case.solution().clone())
            status.append(st)
        runs[name] = (sols, status)
    # step 1 has no step history (no network); then rejected, rejected, paused x3, ...
    R, P = M.GUESS_REJECTED, M.GUESS_PAUSED
    assert runs["linear"][1] == [M.GUESS_USED] * n_steps
    assert runs["learned"][1] == [M.GUESS_USED, R, R, P, P, P, R, R, P, P, P, R]
    for a, b in zip(runs["linear"][0], runs["learned"][0]):
        assert torch.equal(a, b)
