"""The C-ABI library loads, exports every declared entry point and reports
errors correctly, without a GPU (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from pulsar_rocm import native, workloads
from pulsar_rocm.problem import PR_GATE, PR_GRAD_EXACT, PR_PAIRS_SQUARE, COptions, Problem, to_c

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pulsar_rocm.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pr_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_match_binding():
    decl = declared_functions()
    assert decl, "no pr_* declarations found"
    assert sorted(native.EXPORTS) == decl


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_abi_version_and_no_device_here():
    L = native.load()
    assert L.pr_abi_version() == 2
    if not os.path.exists("/dev/kfd"):
        assert native.device_count() == 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU error path")
def test_plan_without_device_fails_cleanly():
    p, u = workloads.four_qubit_ensemble(n_t=4, n_det=1, n_amp=2)
    with pytest.raises(native.RocmError) as e:
        native.Plan([p])
    assert e.value.code == -2  # PR_ERR_DEVICE


def test_validation_runs_before_device_use():
    # a non-Hermitian drift is rejected as PR_ERR_ARG (QuantumSystem.jl Hermiticity check)
    p, u = workloads.four_qubit_ensemble(n_t=4, n_det=1, n_amp=2)
    H0 = np.array(p.H0)
    H0[0, 0, 1] += 1e-6
    p.H0 = H0
    with pytest.raises(native.RocmError) as e:
        native.Plan([p])
    assert e.value.code == -1 and "Hermitian" in str(e.value)


def test_options_validation_runs_before_device_use():
    # segment below PR_SEGMENT_AUTOTUNE (-1), an unknown precision or store_props
    # are PR_ERR_ARG before any device work
    p, u = workloads.four_qubit_ensemble(n_t=4, n_det=1, n_amp=2)
    for kw, what in ((dict(segment=-2), "segment"), (dict(precision=7), "precision"), (dict(store_props=3), "store_props")):
        with pytest.raises(native.RocmError) as e:
            native.Plan([p], **kw)
        assert e.value.code == -1 and what in str(e.value)


def test_non_unitary_target_rejected():
    p, u = workloads.four_qubit_ensemble(n_t=4, n_det=1, n_amp=2)
    p.U_target = 1.01 * np.asarray(p.U_target)
    with pytest.raises(native.RocmError) as e:
        native.Plan([p])
    assert e.value.code == -1 and "unitary" in str(e.value)


def test_injected_plan_failure(monkeypatch):
    monkeypatch.setenv("PULSAR_ROCM_FAULT", "plan")
    p, u = workloads.four_qubit_ensemble(n_t=4, n_det=1, n_amp=2)
    with pytest.raises(native.RocmError) as e:
        native.Plan([p])
    assert e.value.code == -6


def test_injected_eval_failure(monkeypatch):
    monkeypatch.setenv("PULSAR_ROCM_FAULT", "eval")
    L = native.load()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.pr_eval(ctypes.c_void_p(1), None, None, None, None)
    assert rc == -6
    assert "PULSAR_ROCM_FAULT" in native.last_error()


def test_null_arguments_rejected():
    L = native.load()
    h = ctypes.c_void_p()
    assert L.pr_plan_create(None, 1, None, ctypes.byref(h)) == -1
    assert L.pr_eval(None, None, None, None, None) == -1
    assert L.pr_plan_destroy(None) == 0
    assert L.pr_plan_n_params(None) == -1


def test_oversized_dim_unsupported():
    # d > 128 at all; d > 64 outside closed-dynamics gate problems; exact mode at d > 64
    rng = np.random.default_rng(0)
    d = 129
    H = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    H = (H + H.conj().T) / 2
    big = Problem(dim=d, n_ctrl=1, n_t=2, H0=H, Hc=H[None], dt=0.1, objective=PR_GATE,
                  U_target=np.eye(d, dtype=complex))
    with pytest.raises(native.RocmError) as e:
        native.Plan([big])
    assert e.value.code == -5
    p, u = workloads.spin_chain_128(n_t=2, n_members=1)
    with pytest.raises(native.RocmError) as e:
        native.Plan([p], grad_mode=PR_GRAD_EXACT)
    assert e.value.code == -5
    pairs = Problem(dim=80, n_ctrl=1, n_t=2, H0=np.eye(80, dtype=complex), Hc=np.eye(80, dtype=complex)[None],
                    dt=0.1, objective=PR_PAIRS_SQUARE, psi_init=np.eye(80, dtype=complex)[:1],
                    psi_targ=np.eye(80, dtype=complex)[:1])
    with pytest.raises(native.RocmError) as e:
        native.Plan([pairs])
    assert e.value.code == -5


def test_problem_struct_mirrors_match_the_header():
    # pr_problem field order in the header, the ctypes mirror and the Julia mirror
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct \{(.*?)\} pr_problem;", text, flags=re.S).group(1)
    c_fields = re.findall(r"(\w+)\s*;", body)
    from pulsar_rocm.problem import CProblem
    py_fields = [f[0] for f in CProblem._fields_]
    jl = open(os.path.join(ROOT, "pulsar.jl_amd", "julia", "ROCmBackend.jl")).read()
    jl_body = re.search(r"struct PRProblem\n(.*?)\nend", jl, flags=re.S).group(1)
    jl_fields = re.findall(r"^\s*(\w+)::", jl_body, flags=re.M)
    assert c_fields == py_fields == jl_fields, (c_fields, py_fields, jl_fields)


@pytest.mark.parametrize("c_name,py_name,jl_name,jl_file", [
    ("pr_options", "COptions", "PROptions", "ROCmBackend.jl"),
    ("pr_grape_config", "GrapeConfig", "PRGrapeConfig", "rocm_hooks.jl"),
    ("pr_lbfgs_config", "LbfgsConfig", "PRLbfgsConfig", "rocm_hooks.jl"),
    ("pr_mr_config", "MrConfig", "PRMrConfig", "rocm_hooks.jl"),
])
def test_config_struct_mirrors_match_the_header(c_name, py_name, jl_name, jl_file):
    # field order of the option and loop-configuration structs in the header,
    # the ctypes mirrors and the Julia mirrors
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct \{([^{}]*?)\} " + c_name + ";", text, flags=re.S).group(1)
    c_fields = re.findall(r"(\w+)\s*;", body)
    import pulsar_rocm.native as nat
    import pulsar_rocm.problem as prob
    py = getattr(nat, py_name, None) or getattr(prob, py_name)
    py_fields = [f[0] for f in py._fields_]
    jl = open(os.path.join(ROOT, "pulsar.jl_amd", "julia", jl_file)).read()
    jl_body = re.search(r"struct " + jl_name + r"\n(.*?)\nend", jl, flags=re.S).group(1)
    jl_fields = re.findall(r"^\s*(\w+)::", jl_body, flags=re.M)
    assert c_fields == py_fields == jl_fields, (c_fields, py_fields, jl_fields)


def test_julia_constants_match_the_header():
    # every PR_* constant the Julia glue defines has the header's value
    defs = dict((k, int(v)) for k, v in re.findall(r"#define (PR_[A-Z0-9_]+) \(?(-?\d+)\)?", open(HEADER).read()))
    jl = "".join(open(os.path.join(ROOT, "pulsar.jl_amd", "julia", f)).read()
                 for f in ("ROCmBackend.jl", "rocm_hooks.jl"))
    consts = re.findall(r"^const (PR_[A-Z0-9_]+) = Cint\((-?\d+)\)", jl, flags=re.M)
    assert len(consts) >= 14
    for k, v in consts:
        assert defs[k] == int(v), k


@pytest.mark.parametrize("case,code", [("state_lindblad", -5), ("helper_exact", -5), ("helper_state_exact", -5),
                                       ("helper_wide", -5), ("helper_nonunitary", -1), ("modulus_lindblad", -5)])
def test_new_objectives_validate_before_device_use(case, code):
    from pulsar_rocm.problem import PR_GATE_PROPAGATORS, PR_LINDBLAD, PR_STATE, PR_STATE_PROPAGATORS
    d = 80 if case == "helper_wide" else 4
    rng = np.random.default_rng(0)
    H0 = np.diag(rng.standard_normal(d)).astype(complex)
    Hc = np.stack([np.diag(rng.standard_normal(d)).astype(complex)])
    kw = {}
    if case in ("state_lindblad", "modulus_lindblad"):
        from pulsar_rocm.problem import PR_PAIRS_MODULUS
        obj = PR_STATE if case == "state_lindblad" else PR_PAIRS_MODULUS
        p = Problem(dim=d, n_ctrl=1, n_t=3, H0=H0, Hc=Hc, dt=0.1, dynamics=PR_LINDBLAD, objective=obj,
                    psi_init=np.eye(d, dtype=complex)[:1], psi_targ=np.eye(d, dtype=complex)[:1])
    elif case == "helper_state_exact":
        p = Problem(dim=d, n_ctrl=1, n_t=3, H0=H0, Hc=Hc, dt=0.1, objective=PR_STATE_PROPAGATORS,
                    psi_init=np.eye(d, dtype=complex)[:1], psi_targ=np.eye(d, dtype=complex)[:1])
        kw["grad_mode"] = PR_GRAD_EXACT
    else:
        Ut = 1.01 * np.eye(d, dtype=complex) if case == "helper_nonunitary" else np.eye(d, dtype=complex)
        p = Problem(dim=d, n_ctrl=1, n_t=3, H0=H0, Hc=Hc, dt=0.1, objective=PR_GATE_PROPAGATORS, U_target=Ut)
        if case == "helper_exact":
            kw["grad_mode"] = PR_GRAD_EXACT
    with pytest.raises(native.RocmError) as e:
        native.Plan([p], **kw)
    assert e.value.code == code, str(e.value)
