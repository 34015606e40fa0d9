"""The learned initial guess on the GPU: forward and gradients of the BF16
matrix-core kernels against the CPU oracle (oracle/cnn_oracle.py, which
emulates the same bfloat16 rounding points), training on the library's own
FP64 solves, the residual-based fallback, and the measured effect of the
trained network on the SOR sweep count."""
import numpy as np
import pytest

from oracle import cnn_oracle as N
from ttcode_amd import params as P
from ttcode_amd.cnn_init import init_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from ttcode_amd import ttgpu
    d = ttgpu.Device(P.make_params(16, 8, 6))
    yield d
    d.cnn_release()
    d.close()


def _case(seed):
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal((6, 8, 16, 5)).astype(np.float32)
    tgt = rng.standard_normal((6, 8, 16)).astype(np.float32)
    w, b = init_params(seed)
    w = (w * 2.0).astype(np.float32)          # every layer contributes
    b = (rng.standard_normal(b.size) * 0.1).astype(np.float32)
    return x0, tgt, w, b


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle(dev):
    x0, _, w, b = _case(1)
    dev.cnn_set_weights(w, b)
    y = dev.cnn_eval(x0)
    yo, _ = N.forward(x0.astype(np.float64), w, b, bf=True)
    e = _rel(y, yo)
    print(f"forward rel L2 error vs oracle: {e:.2e}")
    # bf16 activations: the oracle and the kernels round at the same points;
    # fp32 summation order and tanhf can flip a rounding now and then
    assert e < 5e-3, e


def test_gradients_match_oracle(dev):
    x0, tgt, w, b = _case(2)
    dev.cnn_set_weights(w, b)
    y, loss, gw, gb = dev.cnn_eval(x0, tgt)
    lo, gwo, gbo, _ = N.loss_grad(x0.astype(np.float64), tgt.astype(np.float64), w, b, bf=True)
    assert abs(loss / lo - 1) < 1e-3, (loss, lo)
    off = 0
    for l in range(N.NL):
        n = 27 * N.CH[l] * N.CH[l + 1]
        e = _rel(gw[off:off + n], gwo[off:off + n])
        print(f"layer {l} weight-gradient rel L2 error {e:.2e}")
        assert e < 2e-2, (l, e)
        off += n
    e = _rel(gb, gbo)
    print(f"bias-gradient rel L2 error {e:.2e}")
    assert e < 2e-2, e


def test_training_on_own_solves_reduces_loss():
    from ttcode_amd import cnn_train as T
    d = T.collect(24, steps_per_config=4, seed=11)
    try:
        assert d.cnn_dataset_size() == 24
        losses = T.train(d, epochs=12, batch=4, lr=3e-3)
        print("epoch losses", losses)
        assert np.all(np.isfinite(losses))
        assert min(losses[-3:]) < 0.5 * losses[0]
    finally:
        d.cnn_release()
        d.close()


def _energy(p, st, guess, weights=None):
    from ttcode_amd import ttgpu
    d = ttgpu.Device(p)
    d.load_state(st)
    d.bounds()
    d.pressg()
    dt = d.schritt()
    d.load_state(st)
    if weights is not None:
        d.cnn_set_weights(*weights[:2])
        d.cnn_set_norm(weights[2])
    d.set_solver(guess=guess)
    es, _ = d.coupled_energy(dt, 2)
    out = {k: d.download(k) for k in ("T", "ER", "delta_ER")}
    d.close()
    return es, out


def test_adversarial_guess_falls_back_to_cold_start():
    """A guess whose residual exceeds the cold start's is rejected; the solve
    is then the cold-start solve, bit for bit."""
    p = P.make_params(32, 64, 32, epsmat=1e-8)
    st = P.generate_state(p, seed=5)
    w, b = init_params(3)
    w = (w * 50.0).astype(np.float32)         # wild network
    b = (b + 5.0).astype(np.float32)
    norm = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1e-2], np.float32)
    es_c, cold = _energy(p, st, 0)
    es_l, lrn = _energy(p, st, 2, (w, b, norm))
    from ttcode_amd import ttgpu
    ttgpu.load_library().tt_cnn_release()
    assert es_l.fallbacks >= 1 and es_l.learned_used == 0
    assert es_l.sor_iter_total == es_c.sor_iter_total
    for k in cold:
        np.testing.assert_array_equal(lrn[k], cold[k], err_msg=k)


def test_rejected_network_is_skipped_then_retried():
    """After 3 first solves in a row reject the learned guess, the network
    runs only on every 16th first solve.  The others start warm or cold,
    whichever has taken less time on recent first solves (moving averages,
    weight 1/4, of the first Newton iteration's SOR phase, by the start the
    solve was given; an untried warm start is tried first; every 16th such
    solve takes the other start).  The policy is replayed here from the
    per-call statistics.  A rejected guess never changes the solve."""
    p = P.make_params(16, 32, 16, epsmat=1e-8)
    st = P.generate_state(p, seed=9)
    w, b = init_params(3)
    w = (w * 50.0).astype(np.float32)
    b = (b + 5.0).astype(np.float32)
    norm = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1e-2], np.float32)
    from ttcode_amd import ttgpu
    d = ttgpu.Device(p)
    d.load_state(st)
    d.bounds()
    d.pressg()
    dt = d.schritt()
    d.load_state(st)
    d.cnn_set_weights(w, b)
    d.cnn_set_norm(norm)
    d.set_solver(guess=2)
    ran, starts = [], []
    avg, picks = [0.0, 0.0], 0
    try:
        for s in range(22):
            es, _ = d.coupled_energy(dt, 2 + s)
            ran.append(1 - es.learned_skipped)
            assert es.learned_used == 0 and es.first_start in (0, 1)
            if es.learned_skipped:
                prefer_warm = avg[1] == 0.0 or (avg[0] > 0.0 and avg[1] <= avg[0])
                picks += 1
                want_warm = (not prefer_warm) if picks % 16 == 0 else prefer_warm
                staged = es.guess_scale_first != 0.0
                assert staged == want_warm, (s, avg, picks)
                if not want_warm:
                    assert es.first_start == 0
            else:
                assert es.first_start == 0  # the wild network's guess is rejected
            starts.append(es.first_start)
            assert es.first_choice == (2 if not es.learned_skipped else (1 if staged else 0))
            k = es.first_choice
            if k < 2:
                t = es.ms_sor_first
                assert t > 0.0
                avg[k] = t if avg[k] == 0.0 else 0.75 * avg[k] + 0.25 * t
    finally:
        d.close()
        ttgpu.load_library().tt_cnn_release()
    # runs on calls 0, 1, 2, then is skipped 15 times, runs on call 18, skipped after
    print("network ran on calls", [i for i, r in enumerate(ran) if r], "starts", starts, "averages", avg)
    assert ran[:3] == [1, 1, 1] and sum(ran[3:18]) == 0 and ran[18] == 1 and sum(ran[19:]) == 0
    assert 1 in starts  # the warm start was tried


def test_trained_guess_effect_on_sweeps():
    """Trains on 200 of the library's own solves at 32x64x32 and reports, on
    held-out configurations, the SOR iterations of the first Newton solve
    with the cold start, the warm start and the learned guess, together with
    the least-residual factor applied to the guess and its residual relative
    to the cold start's.  The learned path must reach the same tolerance and
    agree with the cold-start solution within it; a guess is only ever used
    when its residual is not larger."""
    from ttcode_amd import cnn_train as T
    from ttcode_amd import ttgpu
    d = T.collect(200, steps_per_config=4, seed=1)
    losses = T.train(d, epochs=24, batch=8, lr=3e-3)
    weights = (*d.cnn_get_weights(), d.cnn_get_norm())
    d.close()
    print("training loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.2 * losses[0]
    rng = np.random.default_rng(999)
    rows = []
    for c in range(3):
        p, knobs = T.sample_params(rng)
        st = P.generate_state(p, seed=int(rng.integers(1 << 30)))
        res = {}
        for guess in (0, 1, 2):
            dev = ttgpu.Device(p)
            dev.cnn_set_weights(*weights[:2])
            dev.cnn_set_norm(weights[2])
            dev.load_state(st)
            dev.bounds()
            dev.pressg()
            dt = dev.schritt()
            dev.load_state(st)
            dev.set_solver(guess=guess)
            its, scales, r0s = [], [], []
            for s in range(3):
                e, _ = dev.step(dt, 2 + s)
                en = e.energy
                its.append(en.sor_iter_first)
                scales.append(en.guess_scale_first)
                r0s.append(en.guess_r0_first)
                assert en.guess_r0_first <= 1.0 or en.fallbacks > 0
            res[guess] = (its, scales, r0s, dev.download("T"))
            dev.close()
        for guess in (1, 2):
            err = float(np.max(np.abs(res[guess][3] - res[0][3]) / res[0][3]))
            assert err < 1e-6, (guess, err)
        rows.append((knobs, res))
        print(f"config {c} (a, b, amp) = {np.round(knobs, 3)}")
        for guess, name in ((0, "cold"), (1, "warm"), (2, "learned")):
            its, scales, r0s, _ = res[guess]
            print(f"   {name:8s} first-solve SOR iterations {its}  scale {np.round(scales, 4)}  "
                  f"r0/r0_cold {np.round(r0s, 4)}")
    cold = sum(sum(r[1][0][0]) for r in rows)
    learned = sum(sum(r[1][2][0]) for r in rows)
    warm = sum(sum(r[1][1][0]) for r in rows)
    print(f"total first-solve iterations: cold {cold}, warm {warm}, learned {learned} "
          f"(learned saves {100.0 * (cold - learned) / cold:.1f}%)")
    ttgpu.load_library().tt_cnn_release()
