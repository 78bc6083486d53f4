"""Host-side API mirror without a device: the device registry
(DeviceRegistry.jl:51-151) and the MR context defaults (OptControl.jl:202-259,
733-885)."""
import threading
import warnings

import numpy as np
import pytest

import pulsar_rocm as P
from pulsar_rocm import api, mr
from pulsar_rocm.problem import PR_PAIRS_MODULUS, PR_PAIRS_REAL, PR_PAIRS_SQUARE

SX = np.array([[0, 1], [1, 0]], complex) / 2
SY = np.array([[0, -1j], [1j, 0]], complex) / 2
SZ = np.diag([0.5 + 0j, -0.5])


@pytest.fixture(autouse=True)
def reset_device():
    prev = api._device_default
    yield
    api._device_default = prev
    if hasattr(api._device_local, "device"):
        del api._device_local.device


def test_registry_whitelist_and_unavailable_devices():
    P.set_device("cpu")
    assert P.get_device() == "cpu"
    with pytest.raises(ValueError):
        P.set_device("tpu")
    with pytest.raises(ValueError):
        with P.with_device(":tpu"):
            pass
    # a device whose backend is not loaded warns and keeps the current one
    for dev in ("cuda", ":metal"):
        with pytest.warns(UserWarning):
            P.set_device(dev)
        assert P.get_device() == "cpu"
    if not api.rocm_available():
        with pytest.warns(UserWarning):
            P.set_device("rocm")
        assert P.get_device() == "cpu"


def test_available_devices():
    devs = P.available_devices()
    assert devs[0] == "cpu"
    assert ("rocm" in devs) == api.rocm_available()
    assert set(devs) <= {"cpu", "rocm"}


def test_with_device_restores_nested_and_on_error():
    P.set_device("cpu")
    with P.with_device("rocm"):
        assert P.get_device() == "rocm"
        with P.with_device(":cpu"):
            assert P.get_device() == "cpu"
        assert P.get_device() == "rocm"
    assert P.get_device() == "cpu"
    with pytest.raises(RuntimeError):
        with P.with_device("rocm"):
            raise RuntimeError("inside")
    assert P.get_device() == "cpu"


def test_task_local_override_and_global_default():
    # with_device overrides only the calling thread; other threads see the global default
    seen = {}

    def other():
        seen["dev"] = P.get_device()

    P.set_device("cpu")
    with P.with_device("rocm"):
        t = threading.Thread(target=other)
        t.start()
        t.join()
    assert seen["dev"] == "cpu"


def _ctrl(**kw):
    return mr.MRControl(drifts=[0.3 * SZ], operators=[SX, SY], rho_init=[np.array([1, 0], complex)],
                        rho_targ=[np.array([0, 1], complex)], pulse_dt=np.full(4, 0.1), **kw)


def test_mrcontrol_default_fidelity_is_real():
    # MRControl's keyword default is :real (OptControl.jl:217)
    c = _ctrl()
    assert c.fidelity == "real"
    assert c.problem().objective == PR_PAIRS_REAL
    assert _ctrl(fidelity=":square").problem().objective == PR_PAIRS_SQUARE
    assert _ctrl(fidelity="modulus").problem().objective == PR_PAIRS_MODULUS
    with pytest.raises(ValueError):
        _ctrl(fidelity="dm_uhlmann").problem()


def test_lindblad_mrcontrol_default_and_modulus_rejected():
    # LindbladMRControl defaults to :square (OptControl.jl:818); lindblad_grad_prefactor
    # has :square and :real only (Lindblad.jl:220-232)
    kw = dict(drifts=[0.3 * SZ], operators=[SX, SY], rho_init=[np.array([1, 0], complex)],
              rho_targ=[np.array([0, 1], complex)], pulse_dt=np.full(4, 0.1))
    c = mr.LindbladMRControl(**kw)
    assert c.fidelity == "square" and c.problem().objective == PR_PAIRS_SQUARE
    with pytest.raises(ValueError):
        mr.LindbladMRControl(**kw, fidelity="modulus").problem()
