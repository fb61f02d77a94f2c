"""Engine variants (ostft_engine_variant, ostft_autotune_batch): the radix plan
of the FP32 single-kernel transform and the workgroup shape of the batch
kernels.

Plans: 0 radix-16 passes, remainder radix last; 1 remainder first; 2 radix-8
passes (8 values per thread), remainder last; 3 radix 8, remainder first.

Bars:
  * a radix plan is the transform's arithmetic: every built plan other than 0
    is checked against the oracle at the parity suite's bar (per-frame
    relative L2 <= 1e-5 against the float32 restatement of the reference
    Serial runner; <= 1e-5 against the FP64 pipeline for the new layouts), at
    every window it is built for;
  * a workgroup shape only groups frames: under every (shape, plan) every
    request of a ragged batch is bitwise identical to the same request run
    alone with ostft_run_device on the same context (which uses the same
    plan).
"""
import numpy as np
import pytest

from conftest import frame_errors

import occastft_amd as ostft
from occastft_amd import _native as N

import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-5
EXPS = range(6, 16)                     # window exponents of the batch kernels


def plans_of(exp):
    """Built radix plans of window 2^exp (complex size 2^(exp-1))."""
    lm = exp - 1
    plans = [0] + ([1] if lm % 4 else [])
    if 7 <= lm <= 13:
        plans += [2] + ([3] if lm % 3 else [])
    return plans


def shapes_of(exp, plan=0):
    if plan >= 2:
        return [0]
    return [0, 1] + ([2] if exp <= 12 else []) if exp <= 13 else [0]


PLAN1 = [e for e in EXPS if 1 in plans_of(e)]  # sizes with a remainder radix: all but 9 and 13
OTHER_PLANS = [(e, p) for e in EXPS for p in plans_of(e) if p]


@pytest.fixture(scope="module")
def runner():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    r = ostft.Runner(0)
    yield r
    r.close()


@pytest.fixture(autouse=True)
def reset_variants(runner):
    yield
    for e in EXPS:
        runner.engine_variant(e, 0, 0)
        runner.batch_config(e, blocks_per_cu=-1)


def check(y, r, per, what):
    mx, p50, zabs = frame_errors(y, r, per)
    assert mx <= TOL, f"{what}: max per-frame rel err {mx:.3e} (p50 {p50:.3e})"
    assert zabs <= 1e-6, f"{what}: zero-frame abs {zabs:.3e}"


def bitwise_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_variant_validity(runner):
    for e in EXPS:
        assert runner.engine_variant(e) == (0, 0)
        for plan in range(5):
            if plan not in plans_of(e):
                with pytest.raises(ostft.STFTError) as ei:
                    runner.engine_variant(e, plan=plan)
                assert ei.value.status == N.ERR_UNSUPPORTED
            else:
                assert runner.engine_variant(e, plan=plan)[1] == plan
        runner.engine_variant(e, 0, 0)
        for shape in range(4):
            if shape in shapes_of(e):
                assert runner.engine_variant(e, shape=shape)[0] == shape
            else:
                with pytest.raises(ostft.STFTError) as ei:
                    runner.engine_variant(e, shape=shape)
                assert ei.value.status == N.ERR_UNSUPPORTED
        runner.engine_variant(e, 0, 0)
    for e in (5, 16, 20):
        with pytest.raises(ostft.STFTError) as ei:
            runner.engine_variant(e, plan=0)
        assert ei.value.status == N.ERR_UNSUPPORTED
    # the radix-8 plans have the default shape only
    with pytest.raises(ostft.STFTError) as ei:
        runner.engine_variant(10, shape=1, plan=2)
    assert ei.value.status == N.ERR_UNSUPPORTED
    runner.engine_variant(10, shape=1)
    with pytest.raises(ostft.STFTError):
        runner.engine_variant(10, plan=2)
    runner.engine_variant(10, 0, 0)
    # the persistent form has shape 0 only: selecting it resets the shape
    runner.engine_variant(10, shape=1)
    assert runner.batch_config(10, blocks_per_cu=1) == 1
    assert runner.engine_variant(10)[0] == 0
    # and a shape other than 0 selects one workgroup per tile
    runner.engine_variant(10, shape=2)
    assert runner.batch_config(10) == 0


@pytest.mark.parametrize("exp,plan", OTHER_PLANS)
def test_plan_against_oracle(runner, exp, plan):
    rng = np.random.default_rng(7000 + exp)
    N_ = 1 << exp
    x = rng.uniform(-1, 1, 9 * N_ + 333).astype(np.float32)
    cases = [(0.5, "--hanning_window --remove_dc"), (0.75, "--blackman_window --half_complex_return"),
             (0.0, "--remove_dc")]
    base = {}
    for pl in (0, plan):
        runner.engine_variant(exp, plan=pl)
        for ov, opts in cases:
            y = runner.stft(x, exp, ov, opts)
            rc, r = po.pipeline(x, exp, ov, opts)
            assert rc == 0 and y.shape == r.shape
            check(y, r, N_, f"plan {pl} exp={exp} ov={ov} opts='{opts}'")
            base.setdefault((ov, opts), []).append(y)
        # a new layout against the FP64 pipeline
        opts = "--hanning_window --remove_dc --exact_window --exact_odd_exp --power_return"
        y = runner.stft(x, exp, 0.5, opts)
        g = ostft.geometry(ostft.parse_options(opts, exp, 0.5), len(x))
        t = po.stft64(x, exp, g["hop"], g["frames"], po.WINDOWS["hanning"], 0.0, True, po.OUT_POWER,
                      quirk_odd=False)
        check(y, t, N_ // 2 + 1, f"plan {pl} exp={exp} power vs fp64")
    # the plans are different arithmetic: they agree to rounding, and the
    # plan's kernel did run (its results are not bit for bit plan 0's)
    assert any(not bitwise_equal(a, b) for a, b in base.values()), f"exp {exp}: plan {plan} gave plan 0's bits"


@pytest.mark.parametrize("shape,plan,form", [(0, 1, "dispatch"), (0, 1, "persistent"), (1, 0, "dispatch"),
                                             (1, 1, "dispatch"), (2, 0, "dispatch"), (2, 1, "dispatch"),
                                             (0, 2, "dispatch"), (0, 3, "persistent")])
def test_batch_variant_matches_requests_run_alone(runner, shape, plan, form):
    import torch

    from test_batch_serving import device_inputs, ragged_requests, run_alone

    for e in EXPS:
        pl = plan if plan in plans_of(e) else (plan & 1 if (plan & 1) in plans_of(e) else 0)
        runner.engine_variant(e, shape=shape if shape in shapes_of(e, pl) else 0, plan=pl)
        if form == "persistent":
            runner.batch_config(e, blocks_per_cu=2)
    reqs = ragged_requests(100 + 10 * shape + plan, 48, exps=list(EXPS))
    alone = [run_alone(runner, x, p) for x, p in reqs]
    xs = device_inputs(reqs, misalign=True)
    outs = runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])
    torch.cuda.synchronize()
    for i, (o, a) in enumerate(zip(outs, alone)):
        assert bitwise_equal(o.cpu().numpy(), a), f"request {i}: exp {reqs[i][1].exp} out {reqs[i][1].output}"


@pytest.mark.parametrize("exp,plan", [(6, 0), (6, 1), (7, 0), (7, 1), (8, 1), (9, 0), (10, 0), (10, 2), (11, 1),
                                      (8, 2), (11, 3), (14, 2)])
def test_plan_any_hop(runner, exp, plan):
    """A plan with an explicit geometry: hop 1, a hop past the window (fewer
    frames per tile than the workgroup holds), both N/2+1 layouts, against
    the FP64 pipeline (windows 256 to 2048 write these layouts a wave per
    frame run, windows 64 and 128 with 4 and 2 frames side by side in a
    wave)."""
    from test_mfma import _activate_geometry

    runner.engine_variant(exp, plan=plan)
    N_ = 1 << exp
    rng = np.random.default_rng(exp)
    for hop in (1, N_ // 2 + 3, 3 * N_ + 5):
        frames = 37
        x = rng.uniform(-1, 1, (frames - 1) * hop + N_ // 2).astype(np.float32)
        for out, mode, per in ((N.OUT_POWER, po.OUT_POWER, N_ // 2 + 1),
                               (N.OUT_HALF_SPECTRUM, po.OUT_HALF_SPECTRUM, N_ + 2)):
            p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, out, N.FLAG_REMOVE_DC | N.FLAG_EXACT_WINDOW)
            y = _activate_geometry(runner, x, p, hop, frames)
            t = po.stft64(x, exp, hop, frames, po.WINDOWS["hanning"], 0.0, True, mode, quirk_odd=False)
            check(y, t, per, f"exp {exp} plan {plan} hop {hop} layout {out}")


@pytest.mark.parametrize("like", [None, "--hamming_window --power_return@0.75", "--half_spectrum_return@0"])
def test_autotune_picks_a_variant_and_keeps_results(runner, like):
    import torch

    from test_batch_serving import device_inputs, ragged_requests, run_alone

    for e in (7, 10, 14):
        lp = None
        if like:
            opts, ov = like.split("@")
            lp = ostft.parse_options(opts, 11, float(ov))  # its exp is ignored
        choice = runner.autotune_batch(e, 1 << 23, like=lp)
        assert choice.ms > 0 and choice.blocks_per_cu >= 0
        assert runner.engine_variant(e) == (choice.shape, choice.plan)
        assert runner.batch_config(e) == choice.blocks_per_cu
        assert choice.plan in plans_of(e) and choice.shape in shapes_of(e, choice.plan)
        assert choice.blocks_per_cu == 0 or choice.shape == 0
        reqs = ragged_requests(50 + e, 10, exps=[e])
        alone = [run_alone(runner, x, p) for x, p in reqs]
        xs = device_inputs(reqs)
        outs = runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])
        torch.cuda.synchronize()
        for o, a in zip(outs, alone):
            assert bitwise_equal(o.cpu().numpy(), a)


def test_autotune_rejects_a_bad_like(runner):
    """A template the batch kernels cannot run, or an invalid one, is refused
    before anything is timed, and the context's settings stay."""
    before = runner.engine_variant(10), runner.batch_config(10)
    for bad in (ostft.parse_options("--hanning_window --mfma_fp16", 10, 0.5), ostft.make_params(10, 1.0)):
        with pytest.raises(ostft.STFTError) as ei:
            runner.autotune_batch(10, 1 << 20, like=bad)
        assert ei.value.status == N.ERR_INVALID_ARG
    assert (runner.engine_variant(10), runner.batch_config(10)) == before
