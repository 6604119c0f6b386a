"""Learned warm start end to end on the GPU: ground states generated with the
batched IT solver, an FNO trained on them with the project's own kernels, and
IT from the FNO's prediction against plain MULTIPLE_GAUSS on held-out
instances.  A small network and training set, so the test runs in seconds;
the full-size experiment is gpesolver_amd.warmstart_experiment."""
import numpy as np
import pytest

from gpesolver_amd import fno as F
from gpesolver_amd import warmstart as W

pytestmark = pytest.mark.gpu


def test_warm_start_needs_fewer_it_steps():
    rng = np.random.default_rng(7)
    insts = W.sample_instances(rng, 40)
    train, test = insts[:32], insts[32:]
    X, T, S = W.generate(train, tol=1e-6, max_steps=5000, log=lambda m: print(m, flush=True))
    model = F.Fno(*W.GRID, width=16, layers=2, hidden=32, modes=8, max_batch=8, seed=3)
    hist = W.train(model, X, T, epochs=30, lr=2e-3, batch=8, log=lambda m: None)
    print("train mse first / last", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.5 * hist[0]
    res = W.compare(model, test, tol=1e-6, max_steps=5000)
    model.close()
    print({k: res[k] for k in ("plain_steps_mean", "warm_steps_mean", "plain_wall_s", "warm_wall_s",
                               "fallbacks", "energy_rel_diff_median")}, flush=True)
    assert res["fallbacks"] == 0
    assert res["warm_steps_mean"] < res["plain_steps_mean"]
    # the warm run reaches the plain run's energy (or a lower one)
    pe, we = np.array(res["plain_e_per_atom"]), np.array(res["warm_e_per_atom"])
    assert np.all(we <= pe + 1e-4 * np.abs(pe))


def test_fallback_on_nan_and_zero_warm_state():
    """A NaN and a zero warm state, given to gpe_load_buffer (which cannot
    normalise them), make the batch restart those instances from their
    MULTIPLE_GAUSS fallback at the first check: they then reach the plain
    run's E/N exactly, check_every steps later.  An instance with a good
    warm state (here its own plain state) does not fall back."""
    rng = np.random.default_rng(11)
    insts = W.sample_instances(rng, 3)
    plain = [W.build(d)[2] for d in insts]
    s_p, e_p, _, _ = W.run_batch(insts, plain, tol=1e-6, max_steps=40000)
    bad = [np.full(plain[0].shape, np.nan, dtype=np.complex128), np.zeros(plain[1].shape, dtype=np.complex128),
           plain[2]]
    s_w, e_w, _, _, taken = W.run_batch(insts, bad, tol=1e-6, max_steps=40000, fallback=True, with_taken=True)
    assert s_p.max() < 39000  # all converge before the cap
    print("plain", s_p, e_p, "fallback run", s_w, e_w, taken, flush=True)
    assert taken.tolist() == [True, True, False]
    np.testing.assert_array_equal(e_w, e_p)
    np.testing.assert_array_equal(s_w, s_p + np.array([100, 100, 0]))


def test_fallback_on_energy_rise():
    """An instance whose explicit imaginary-time map is unstable (dtau 6.3x
    the reference value; the kinetic stencil alone has dtau lambda_max ~2.1
    at 120 / 600 nm) grows a mode from round-off within the first 100 steps
    and would then meet the |dE/N| test on it (E/N ~6e4 times the ground
    state's); its dtau mu >= 1 sends it to its fallback.  The same instance at
    the reference dtau lowers E/N at every check and keeps its warm state."""
    import dataclasses

    from gpesolver_amd import hip

    rng = np.random.default_rng(11)
    d = W.sample_instances(rng, 1)[0]
    p, V, plain = W.build(d)
    p_bad = dataclasses.replace(p, imag_time_dt=p.imag_time_dt * 6.3)
    b = hip.Batch([p_bad.to_gpe(), p.to_gpe()], hip.GPE_FP32)
    try:
        for q, dev in enumerate(b.instances):
            dev.init_potential(d["pot"], p)
            dev.load_buffer(plain)
            b.set_fallback(q, plain)
        steps, e = b.it_until(1500, 100, 1e-6)
        taken = b.fallbacks_taken()
    finally:
        b.close()
    print("steps", steps, "E/N", e, "taken", taken, flush=True)
    assert taken.tolist() == [True, False]


@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "fp64"])
def test_device_warm_start_equals_host_path(prec):
    """The warm start on the device (gpe_warm_features from the instance's
    potential, gpe_fno_forward_dev, gpe_load_amplitude with and without the
    mirror average) gives the host path's state bit for bit: features of the
    exported potential, forward pass from host arrays, warm_state fed to
    gpe_load_buffer.  The norm returned is the loaded state's dV sum |psi|^2."""
    from gpesolver_amd import hip

    precision = (hip.GPE_FP32, hip.GPE_FP64)[prec]
    rng = np.random.default_rng(21)
    model = F.Fno(*W.GRID, width=8, layers=1, hidden=8, modes=4, max_batch=2, seed=5)
    insts = W.sample_instances(rng, 2)
    w0 = model.get_params()
    try:
        for d in insts:
            p, V, _ = W.build(d)
            # shift the output bias (the last parameter) so that about three
            # quarters of this instance's prediction is positive: the clip to
            # zero and the mirror average both act
            model.set_params(w0)
            y0 = model.forward(np.stack([F.features(p, V, d["pot"])] * 2))[0]
            w = w0.copy()
            w[-1] -= np.float32(np.quantile(y0, 0.25))
            model.set_params(w)
            dev = hip.Device(p.to_gpe(), precision)
            ref = hip.Device(p.to_gpe(), precision)
            try:
                dev.init_potential(d["pot"], p)
                dev.warm_features(model.input_ptr(1), F.CIN, F.V0, F.feature_consts(p, d["pot"]))
                model.forward_device(2)
                x = F.features(p, dev.export_pote(), d["pot"])
                for mirror in (True, False):
                    norm = dev.load_amplitude(model.output_ptr(1), mirror)
                    got = dev.export_psi()
                    y = model.forward(np.stack([x, x]))[1]
                    assert (y > 0).mean() > 0.3 and (y < 0).any(), (float(np.min(y)), float(np.max(y)))
                    psi = F.warm_state(y, W.GRID, symmetric=mirror)
                    ref.load_buffer(psi)
                    np.testing.assert_array_equal(got, ref.export_psi())
                    dv = p.dx * p.dy * p.dz
                    assert abs(norm - dv * np.sum(np.abs(psi) ** 2)) < 1e-5 * norm
                    # the device input was overwritten by the host forward pass: rebuild it
                    dev.warm_features(model.input_ptr(1), F.CIN, F.V0, F.feature_consts(p, d["pot"]))
                    model.forward_device(2)
            finally:
                dev.close()
                ref.close()
    finally:
        model.close()
