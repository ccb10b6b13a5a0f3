"""Inputs the reference would crash or hang on are rejected by the host with
RMS_ERR_INVALID_ARG and a message, before any kernel runs, and the context
stays usable afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INVALID = 1  # RMS_ERR_INVALID_ARG


def _err(rms, fn, text):
    capi = __import__(rms.__name__ + "._capi", fromlist=["RmsError"])
    with pytest.raises(capi.RmsError) as ei:
        fn()
    assert ei.value.status == INVALID, str(ei.value)
    assert text in str(ei.value), str(ei.value)


def test_invalid_inputs_rejected(rms, engine):
    print("case invalid inputs", flush=True)
    ones = np.ones((4, 4, 4), np.uint8)
    ok = dict(geom=rms.Geometry(ones, 0.1, [0.0, 2.0]), nwalk=64, seed=1)
    # no walkable voxel: the reference's initialisation loop would never end
    _err(rms, lambda: engine.run([dict(ok, geom=rms.Geometry(np.zeros((3, 3, 3), np.uint8), 0.1, [0.0, 2.0]))],
                                 2e-4, 100, 10), "no non-null voxel")
    # label beyond the class table
    _err(rms, lambda: engine.run([dict(ok, geom=rms.Geometry(ones * 2, 0.1, [0.0, 2.0]))], 2e-4, 100, 10),
         "label value out of range")
    # a step of one voxel or more (main.cu:210-214 assumes less)
    _err(rms, lambda: engine.run([dict(ok, geom=rms.Geometry(ones, 0.1, [0.0, 9.0]))], 2e-4, 100, 10), "step size")
    _err(rms, lambda: engine.run([dict(ok, nwalk=0)], 2e-4, 100, 10), "nwalk")
    _err(rms, lambda: engine.run([ok], 2e-4, 5, 10), "TN must be >= timepoints")
    _err(rms, lambda: engine.run([ok], 2e-4, 100, 10, block_threads=100), "block_threads")
    g = rms.Geometry(ones, 0.1, [0.0, 2.0])
    g.periodic = 8
    _err(rms, lambda: engine.run([dict(ok, geom=g)], 2e-4, 100, 10), "periodic")
    g = rms.Geometry(ones, 0.1, [0.0, 2.0])
    g.storage = "tiles"
    with pytest.raises(KeyError):
        g.to_c()
    # adjoint options: an unknown D estimator, a step of the central
    # differences outside (0, 0.5)
    plan = engine.plan([ok], 2e-4, 100, 10)
    try:
        cot = [np.ones((10, rms.nval(2, 0)))]
        _err(rms, lambda: plan.adjoint(cot, None, fd_step=0.7), "fd_rel_step")
        _err(rms, lambda: plan.adjoint(cot, None, fd_step=-0.1), "fd_rel_step")
        capi = __import__(rms.__name__ + "._capi", fromlist=["D_ESTIMATOR"])
        capi.D_ESTIMATOR["bogus"] = 9
        try:
            _err(rms, lambda: plan.adjoint(cot, None, estimator="bogus"), "d_estimator")
        finally:
            del capi.D_ESTIMATOR["bogus"]
        assert np.isfinite(plan.adjoint(cot, None)[0]["grad"][1])
    finally:
        plan.close()
    # the context still runs a valid job
    r = engine.run([ok], 2e-4, 100, 10)[0]
    assert r["NPar_count"].sum(1).tolist() == [64.0] * 10
