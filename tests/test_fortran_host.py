"""The Fortran host layer (ferebus-v7_amd/src) without a GPU.

Builds tests/fortran/test_host.f90 with amdflang against the host modules and
checks what it reads and computes against the Python front end, numpy and the
oracle:
* ferebus.toml of the reference's examples (the config surface, with the
  reference's defaults, src/config/*);
* kernel-expression programs (src/interpreter/ grammar) equal to the Python
  compiler's;
* training-set CSV reading (src/csv.f90 header convention);
* the four mean functions (src/mean/) equal to the oracle's.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from ferebus_amd import KernelConfig, KernelExpressionError, compile_kernel

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "ferebus-v7_amd"
EX = REPO / "tests" / "golden" / "examples"
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
SRCS = ["src/ferebus_hip.f90", "src/mean/fb_mean.f90", "src/kernels/fb_kernel_expr.f90", "src/io/fb_config.f90"]


@pytest.fixture(scope="module")
def host_exe(tmp_path_factory):
    if not Path(FLANG).exists():
        pytest.skip("amdflang not available")
    lib = PKG / "lib" / "libferebus_hip.so"
    if not lib.exists():
        pytest.skip("libferebus_hip.so not built")
    out = tmp_path_factory.mktemp("fhost")
    exe = out / "test_host"
    cmd = [FLANG, "-O1", "-J", str(out), "-o", str(exe)] + [str(PKG / s) for s in SRCS] + \
          [str(REPO / "tests" / "fortran" / "test_host.f90"), "-L" + str(PKG / "lib"), "-lferebus_hip",
           "-Wl,-rpath," + str(PKG / "lib")]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def run(exe, toml):
    out = subprocess.run([str(exe), str(EX / toml), str(EX / "WATER_O1_TRAINING_SET.csv")],
                         check=True, capture_output=True, text=True).stdout
    return [l.strip() for l in out.splitlines()]


def field(lines, key):
    for l in lines:
        if l.startswith(key + "="):
            return l[len(key) + 1:].strip()
    raise KeyError(key)


def test_water_config(host_exe):
    lines = run(host_exe, "WATER_ferebus.toml")
    assert field(lines, "system.name") == "WATER"
    assert field(lines, "system.natoms") == "3"
    assert field(lines, "system.nfeats") == "3"          # 3*natoms - 6
    assert field(lines, "atom1.name") == "O1"
    assert field(lines, "atom1.alf").split() == ["1", "2", "3"]
    assert field(lines, "model.mean") == "constant"
    assert field(lines, "model.kernel") == "k1*k2"
    assert field(lines, "model.likelihood") == "marginal"  # default
    assert float(field(lines, "model.noise")) == 1e-10     # default
    assert float(field(lines, "optimiser.search_max")) == 1.0
    assert field(lines, "pso.swarm_size") == "20"
    assert field(lines, "pso.iterations") == "1"
    assert float(field(lines, "pso.inertia_weight")) == 0.789
    assert field(lines, "pso.stopping") == "none"
    assert field(lines, "kernel.k1") == "rbf"
    assert field(lines, "kernel.k2") == "periodic dims 3"


def test_methanol_config(host_exe):
    lines = run(host_exe, "METHANOL_ferebus.toml")
    assert field(lines, "system.nfeats") == "12"
    assert field(lines, "atom1.alf").split() == ["1", "3", "2"]
    assert float(field(lines, "optimiser.search_max")) == 3.0
    assert field(lines, "pso.stopping") == "relative_change"
    assert float(field(lines, "pso.tolerance")) == 1e-8
    assert field(lines, "pso.stall_iterations") == "50"
    assert float(field(lines, "pso.inertia_weight")) == 0.729
    assert field(lines, "notes") == "2"
    assert field(lines, "note.basis-set") == "6-31+g(d,p)"


def test_kernel_programs_match_python(host_exe):
    lines = run(host_exe, "WATER_ferebus.toml")
    ks = {"k1": KernelConfig("k1", "rbf", [1, 2]), "k2": KernelConfig("k2", "periodic", [3])}
    for expr in ("k1+k2", "k1*k2*k1", "(k1+k2)*k1", "2*k1", "k1+k2+k1", "k1-k2"):
        fl = [l for l in lines if l.startswith(f"expr {expr} ")]
        try:
            p = compile_kernel(expr, ks, 3)
        except KernelExpressionError:
            assert any(" error " in l for l in fl), expr
            continue
        types = [int(v) for v in next(l for l in fl if " types " in l).split(" types ")[1].split()]
        ndims = [int(v) for v in next(l for l in fl if " ndims " in l).split(" ndims ")[1].split()]
        name = next(l for l in fl if " name " in l).split(" name ")[1].strip()
        npar = int(next(l for l in fl if " nparams " in l).split(" nparams ")[1])
        assert types == p.node_type and ndims == p.node_ndim, expr
        assert name == p.composition and npar == p.n_params, expr


def test_csv_and_means(host_exe):
    lines = run(host_exe, "WATER_ferebus.toml")
    path = EX / "WATER_O1_TRAINING_SET.csv"
    head = path.read_text().splitlines()[0].split(",")
    raw = np.loadtxt(path, delimiter=",", skiprows=1)
    x = raw[:, [head.index("f1"), head.index("f2"), head.index("f3")]]
    y = raw[:, head.index("iqa")]
    assert field(lines, "csv.shape").split() == [str(len(y)), "3", "26"]
    assert field(lines, "csv.labels").split() == ["q00", "iqa"]
    np.testing.assert_allclose([float(v) for v in field(lines, "csv.x1").split()], x[0], rtol=1e-15)
    assert float(field(lines, "csv.ylast1")) == y[0]
    for kind, name in enumerate(("zero", "constant", "linear", "quadratic")):
        m, *_ = oracle.mean(name, x, y)
        got = [float(v) for v in field(lines, f"mean{kind}").split()]
        np.testing.assert_allclose(got, [m[0], m[1], m.sum()], rtol=1e-11, atol=1e-12)


def _run_toml(exe, tmp_path, pso_block):
    base = (EX / "WATER_ferebus.toml").read_text()
    text = base + "\n" + pso_block
    path = tmp_path / "ferebus.toml"
    path.write_text(text)
    return subprocess.run([str(exe), str(path), str(EX / "WATER_O1_TRAINING_SET.csv")],
                          capture_output=True, text=True)


def test_updater_keys_sanitised(host_exe, tmp_path):
    """swarm_updater / parameter_updater / stopping_criteria are sanitised as
    the reference does (spaces to underscores, lower case, str.f90:85-98),
    aliases map to their updater, and the adaptive table is looked up under
    the given name, then [adaptive] (config/optimiser/pso/pso.f90:124-139)."""
    base = (EX / "WATER_ferebus.toml").read_text()
    assert "[optimiser.pso]" in base and "stopping_criteria" not in base
    out = _run_toml(host_exe, tmp_path, "")
    assert out.returncode == 0, out.stderr
    lines = [l.strip() for l in out.stdout.splitlines()]
    assert field(lines, "pso.swarm_updater") == "global"
    assert field(lines, "pso.parameter_updater") == "constant"
    assert float(field(lines, "pso.delta")) == 1e-7
    # keys of [optimiser.pso] after the file's own section: re-open it
    text = base.replace("[optimiser.pso]", "[optimiser.pso]\nswarm_updater = \"Forced\"\n"
                        "parameter_updater = \"UAPSO\"\nstopping_criteria = \"Relative Change\"", 1)
    text += ("\n[optimiser.pso.forced]\ndelta = 1e-3\n[optimiser.pso.adaptive]\nmax_inertia_weight = 0.9\n"
             "n_social_learning_rate = 7\nsuccess_threshold = 0.25\n"
             "[optimiser.pso.relative_change]\ntolerance = 1e-5\nstall_iterations = 9\n")
    (tmp_path / "f.toml").write_text(text)
    out = subprocess.run([str(host_exe), str(tmp_path / "f.toml"), str(EX / "WATER_O1_TRAINING_SET.csv")],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    lines = [l.strip() for l in out.stdout.splitlines()]
    assert field(lines, "pso.swarm_updater") == "forced"
    assert float(field(lines, "pso.delta")) == 1e-3
    assert field(lines, "pso.parameter_updater") == "unified_adaptive"
    assert field(lines, "pso.stopping") == "relative_change"
    assert float(field(lines, "pso.tolerance")) == 1e-5 and field(lines, "pso.stall_iterations") == "9"
    assert [float(v) for v in field(lines, "pso.adaptive_max").split()] == [0.9, 4.0, 4.0]
    assert field(lines, "pso.adaptive_n").split() == ["10", "10", "7"]
    assert [float(v) for v in field(lines, "pso.adaptive_steps").split()] == [0.01, 0.01, 0.25]


@pytest.mark.parametrize("key,value,needle", [
    ("stopping_criteria", "potential_based", "no implementation"),
    ("stopping_criteria", "Sign Test", "no implementation"),
    ("stopping_criteria", "relative-change", "unknown stopping_criteria"),
    ("swarm_updater", "local", "unknown swarm_updater"),
    ("parameter_updater", "adaptive", "unknown parameter_updater"),
])
def test_unknown_names_fail_closed(host_exe, tmp_path, key, value, needle):
    base = (EX / "WATER_ferebus.toml").read_text()
    text = base.replace("[optimiser.pso]", f"[optimiser.pso]\n{key} = \"{value}\"", 1)
    (tmp_path / "bad.toml").write_text(text)
    out = subprocess.run([str(host_exe), str(tmp_path / "bad.toml"), str(EX / "WATER_O1_TRAINING_SET.csv")],
                         capture_output=True, text=True)
    assert out.returncode != 0
    assert needle in out.stderr + out.stdout
