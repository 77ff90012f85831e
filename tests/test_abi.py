"""C ABI and host-side checks that need no GPU.

* libferebus_hip.so loads and exports every entry point include/ferebus_hip.h
  declares (no compute call is made);
* the kernel-expression front end follows the reference grammar
  (src/interpreter/parser.f90:102-197, ast.f90:173-271);
* the host mean functions agree with the oracle's (src/mean/).
"""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from ferebus_amd import KernelConfig, KernelExpressionError, compile_kernel, fit_mean
from ferebus_amd import _abi
from ferebus_amd.kernel_expr import (LEAF_CONSTANT, LEAF_LINEAR, LEAF_PERIODIC, LEAF_RBF,
                                     LEAF_RBF_CYCLIC, OP_PROD, OP_SUM)

REPO = Path(__file__).resolve().parents[1]
HEADER = REPO / "include" / "ferebus_hip.h"


def declared_entry_points():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fb_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_known_entry_points():
    assert declared_entry_points() == sorted(_abi.EXPORTS)


def test_fortran_module_binds_every_entry_point():
    """The host's binding module (the reference-side binding a maintainer
    adds) has a bind(c) interface for every declared entry point, and the
    binding excerpt in INTEGRATION.md names each one."""
    module = (REPO / "ferebus-v7_amd" / "src" / "ferebus_hip.f90").read_text()
    bound = set(re.findall(r'bind\(c,\s*name="(fb_[a-z_0-9]+)"\)', module))
    assert set(declared_entry_points()) <= bound, sorted(set(declared_entry_points()) - bound)
    integration = (REPO / "INTEGRATION.md").read_text()
    integration = re.sub(r"/\s*\n\s*!\s*", "/ ", integration)   # Fortran comment continuation
    named = set(re.findall(r"\b(fb_[a-z_0-9]+)", integration))
    # the excerpt abbreviates families ("fb_pso_create / init / retry"):
    # expand "fb_<family>_<a> / <b>" into fb_<family>_<b>
    for fam, rest in re.findall(r"\b(fb_[a-z]+)_[a-z_]+((?:\s*/\s*[a-z_]+)+)", integration):
        named |= {f"{fam}_{w.strip()}" for w in rest.split("/") if w.strip()}
    missing = set(declared_entry_points()) - named
    assert not missing, sorted(missing)


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    for name in declared_entry_points():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.library_path())],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fb_[a-z_0-9]+)\b", out))
    assert set(declared_entry_points()) <= exported


def test_library_is_gfx950_only():
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", str(_abi.library_path())], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    archs = set(re.findall(r"gfx\d+", out.stdout))
    assert archs == {"gfx950"}, out.stdout


KS = {
    "k1": KernelConfig("k1", "rbf", [1, 2]),
    "k2": KernelConfig("k2", "periodic", [3]),
    "k3": KernelConfig("k3", "linear", [1]),
    "kc": KernelConfig("kc", "rbf-cyclic"),
    "ka": KernelConfig("ka", "rbf"),
}


@pytest.mark.parametrize("expr,types,ndims", [
    ("k1", [LEAF_RBF], [2]),
    ("k1+k2", [LEAF_RBF, LEAF_PERIODIC, OP_SUM], [2, 1, 0]),
    ("k1*k2", [LEAF_RBF, LEAF_PERIODIC, OP_PROD], [2, 1, 0]),
    # '*' is right recursive: k1*k2*k3 = k1*(k2*k3)
    ("k1*k2*k3", [LEAF_RBF, LEAF_PERIODIC, LEAF_LINEAR, OP_PROD, OP_PROD], [2, 1, 1, 0, 0]),
    # '+' is left associative: k1+k2+k3 = (k1+k2)+k3
    ("k1+k2+k3", [LEAF_RBF, LEAF_PERIODIC, OP_SUM, LEAF_LINEAR, OP_SUM], [2, 1, 0, 1, 0]),
    ("(k1+k2)*k3", [LEAF_RBF, LEAF_PERIODIC, OP_SUM, LEAF_LINEAR, OP_PROD], [2, 1, 0, 1, 0]),
    ("k1+k2*k3", [LEAF_RBF, LEAF_PERIODIC, LEAF_LINEAR, OP_PROD, OP_SUM], [2, 1, 1, 0, 0]),
    ("2*k1", [LEAF_CONSTANT, LEAF_RBF, OP_PROD], [0, 2, 0]),
    ("+k1", [LEAF_RBF], [2]),
    ("kc", [LEAF_RBF_CYCLIC], [4]),
    ("ka", [LEAF_RBF], [4]),
])
def test_kernel_expression_programs(expr, types, ndims):
    p = compile_kernel(expr, KS, 4)
    assert p.node_type == types
    assert p.node_ndim == ndims


def test_parameter_count_and_names():
    p = compile_kernel("k1+k2", KS, 4)
    assert p.n_params == 3 and p.active_dims == [1, 2, 3]
    assert p.composition == "(k1+k2)"
    p = compile_kernel("2*k1", KS, 4)
    assert p.n_params == 3 and p.leaf_names == ["c", "k1"]


@pytest.mark.parametrize("expr", ["k1-k2", "k1/k2", "-k1", "k9", "k1+", "(k1", "k1 k2"])
def test_kernel_expression_errors(expr):
    with pytest.raises(KernelExpressionError):
        compile_kernel(expr, KS, 4)


def test_cyclic_needs_every_feature():
    with pytest.raises(KernelExpressionError):
        compile_kernel("k", {"k": KernelConfig("k", "rbf-cyclic", [1, 2])}, 4)


@pytest.mark.parametrize("kind", ["zero", "constant", "linear", "quadratic"])
def test_host_means_match_oracle(kind):
    rng = np.random.default_rng(5)
    x = rng.uniform(1.6, 2.9, (80, 5))
    y = np.sin(x[:, 0]) + 0.2 * x[:, 3] ** 2 + 0.01 * rng.normal(size=80)
    mf = fit_mean(kind, x, y)
    om, beta, xmin, ymin, info = oracle.mean(kind, x, y)
    np.testing.assert_allclose(mf.value, om, rtol=1e-11, atol=1e-13)
    if kind in ("linear", "quadratic"):
        np.testing.assert_allclose(mf.beta, beta, rtol=1e-10)
        assert mf.ymin == ymin and np.array_equal(mf.xmin, xmin)
