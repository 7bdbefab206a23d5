"""The learned warm start on the GPU, end to end, on the obstacle-channel
workload (256 x 128 D2Q9, 1-4 seeded circles/rectangles, Re 20..200, NEEM
inlet u = 0.05 and outlet, bounce-back walls).

Steady training fields come from the batched FP32 solver; the network is
trained by the project's own HIP kernels (bf16 MFMA forward and backward,
Adam); on unseen channels the plain solver then runs from the predicted
field and, for comparison, from the plain start it replaces -- rho = 1 and
the inlet velocity in every fluid cell, the stable form of the reference's
initialCondition (src/lbm/solver.cpp:267-295, rho = 1, u = 0 plus the
boundary values), whose impulsive inlet makes BGK blow up near tau = 1/2.
Both runs go to the relative residual 1e-6 (||u^n - u^(n-1)|| / ||u^n||,
FP64 sums) and are compared with FP64 solutions converged to 1e-10.

The training set here is small (a few minutes of the full workload would
not fit a test), so the gain is below the full-size figure; what the test
pins is that the trained network saves solver steps on every unseen
channel and that the solver, judging the prediction, converges to the
FP64 solution: at the 1e-6 stop a solve's velocity is ~1e-3 from it
whichever start it had (the stop, not the start, sets that distance), so
the warm-started solve is carried on to 2e-7 and must then be within 1e-3
(relative L2 of u).
"""
import numpy as np
import pytest

from lbm_amd import surrogate as S
from lbm_amd import warmstart as WS

pytestmark = pytest.mark.gpu

NTRAIN = 384
NTEST = 8
EPOCHS = 30


@pytest.mark.timeout(900)
def test_warm_start_saves_steps_and_converges_to_fp64_field():
    print(f"warm start: {NTRAIN} training channels (seed 11), {NTEST} unseen (seed 12)", flush=True)
    doms, x = WS.make_set(NTRAIN, seed=11)
    steps, _, wall, _, fields = WS.solve(doms, 1e-5, 200000)
    ok = np.nonzero(steps > 0)[0]
    print(f"training fields: {ok.size}/{NTRAIN} converged to 1e-5 in {wall:.1f} s", flush=True)
    assert ok.size >= NTRAIN * 3 // 4
    t, m, norm = WS.targets([doms[i] for i in ok], [fields[i] for i in ok])
    xa, ta, ma = WS.mirror(x[ok], t, m)
    sur = S.Surrogate(batch=16, height=WS.NY, width=WS.NX, seed=7)
    losses, train_s = WS.train(sur, xa, ta, ma, EPOCHS, 5e-4, 1e-5)
    print(f"training: loss {losses[0]:.3e} -> {losses[-1]:.3e} in {train_s:.1f} s", flush=True)
    assert losses[-1] < 0.5 * losses[0]

    tdoms, tx = WS.make_set(NTEST, seed=12)
    pred, _ = WS.predict_fields(sur, tx, tdoms, norm)
    sur.close()
    cold, warm = [], []
    for i, d in enumerate(tdoms):
        sc, _, wc, _, fc = WS.solve([d], 1e-6, 400000)
        sw, _, ww, _, fw = WS.solve([d], 1e-6, 400000, fields=[pred[i]])
        cold.append((int(sc[0]), wc, fc[0]))
        warm.append((int(sw[0]), ww, fw[0]))
        print(f"channel {i} (Re {d['re']:.0f}): plain start {sc[0]} steps, warm start {sw[0]} steps", flush=True)
    ref_steps, _, _, _, ref = WS.solve(tdoms, 1e-10, 1600000, storage="f64", fields=[c[2] for c in cold])
    assert (ref_steps > 0).all(), ref_steps
    ratios = []
    for i, d in enumerate(tdoms):
        assert cold[i][0] > 0 and warm[i][0] > 0, (i, cold[i][0], warm[i][0])
        e_pred = WS.rel_l2(pred[i][1], ref[i][1], d["solid"])
        e_cold = WS.rel_l2(cold[i][2][1], ref[i][1], d["solid"])
        e_warm = WS.rel_l2(warm[i][2][1], ref[i][1], d["solid"])
        ratios.append(cold[i][0] / warm[i][0])
        print(f"channel {i}: prediction error {e_pred:.2e}, u after the solve vs FP64: plain {e_cold:.2e}, "
              f"warm {e_warm:.2e}; step ratio {ratios[-1]:.2f}", flush=True)
        # the solver judges the prediction: carried on from its 1e-6 state to
        # 2e-7, the warm-started solve is within the 1e-3 velocity bar
        s2, _, _, _, f2 = WS.solve([d], 2e-7, 400000, fields=[warm[i][2]])
        e_tight = WS.rel_l2(f2[0][1], ref[i][1], d["solid"])
        print(f"channel {i}: warm start carried on to 2e-7 in {int(s2[0])} more steps: u vs FP64 {e_tight:.2e}",
              flush=True)
        assert s2[0] > 0 and e_tight <= 1e-3
        # the prediction is a real head start, far closer than the plain start
        assert e_pred < 0.5
        assert warm[i][0] < cold[i][0]
    print(f"steps saved by the trained network: median ratio {np.median(ratios):.2f}, "
          f"min {min(ratios):.2f}", flush=True)
    assert np.median(ratios) > 1.25


MODEL = WS.SHIPPED_MODEL


@pytest.mark.timeout(600)
def test_shipped_network_saves_steps_on_unseen_channels():
    """The network that ships with the package (trained with these kernels
    on 8192 channels and their mirror images for 240 epochs), on 6 unseen
    channels of another seed: the median step ratio (plain / warm) is above
    1.5 (the small network trained inside the test above reaches ~1.6 on
    its own channels), no warm start costs more than 1.3x the plain start's
    steps (at low Re the plain start is already close and the slow viscous
    modes set the count: one Re 23 channel of this set takes 13,700 steps
    from the prediction against 11,900 from the plain start), and every
    warm-started solve carried on to 2e-7 lands within 1e-3 of the FP64
    field."""
    tdoms, tx = WS.make_set(6, seed=21)
    sur = S.Surrogate(batch=16, height=WS.NY, width=WS.NX, seed=7)
    norm = WS.load_model(sur, MODEL)
    pred, _ = WS.predict_fields(sur, tx, tdoms, norm)
    sur.close()
    ratios = []
    warm_fields, cold_fields = [], []
    for i, d in enumerate(tdoms):
        sc, _, _, _, fc = WS.solve([d], 1e-6, 400000)
        sw, _, _, _, fw = WS.solve([d], 1e-6, 400000, fields=[pred[i]])
        assert sc[0] > 0 and sw[0] > 0, (i, sc[0], sw[0])
        ratios.append(sc[0] / sw[0])
        cold_fields.append(fc[0])
        warm_fields.append(fw[0])
        print(f"channel {i} (Re {d['re']:.0f}): plain start {sc[0]} steps, shipped-network start {sw[0]} steps",
              flush=True)
        assert sw[0] <= 1.3 * sc[0]
    ref_steps, _, _, _, ref = WS.solve(tdoms, 1e-10, 1600000, storage="f64", fields=cold_fields)
    assert (ref_steps > 0).all(), ref_steps
    for i, d in enumerate(tdoms):
        e_pred = WS.rel_l2(pred[i][1], ref[i][1], d["solid"])
        s2, _, _, _, f2 = WS.solve([d], 2e-7, 400000, fields=[warm_fields[i]])
        e_tight = WS.rel_l2(f2[0][1], ref[i][1], d["solid"])
        print(f"channel {i}: prediction error {e_pred:.2e}, warm start carried on to 2e-7: u vs FP64 {e_tight:.2e}",
              flush=True)
        assert e_pred < 0.2 and s2[0] > 0 and e_tight <= 1e-3
    print(f"shipped network: median step ratio {np.median(ratios):.2f}, min {min(ratios):.2f}", flush=True)
    assert np.median(ratios) > 1.5
