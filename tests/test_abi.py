"""The C ABI without a GPU: the library loads, exports every entry point the
header declares, its ctypes mirror has the C layout, and calls that need no
device behave (argument errors first, then MPM_E_NO_DEVICE where no GPU is
visible).  No compute call is made here."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

import code_mpm_auxetics_amd._abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpm_auxetics.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w ]+?\**\s*\b(mpm_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert _declared() == sorted(A.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    if shutil.which("nm"):
        out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True).stdout
        syms = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
        for name in _declared():
            assert name in syms, name


def test_abi_version_matches_header():
    v = int(re.search(r"#define MPM_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert A.lib().mpm_abi_version() == v


def test_struct_layout_matches_c(tmp_path):
    """sizeof / offsetof of mpm_config and mpm_points from a C compile of the
    header equal the ctypes mirror's."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    fields_cfg = [f[0] for f in A.MpmConfig._fields_]
    fields_pts = [f[0] for f in A.MpmPoints._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpm_auxetics.h"', "int main(void){"]
    src.append('printf("%zu %zu\\n", sizeof(mpm_config), sizeof(mpm_points));')
    for f in fields_cfg:
        src.append(f'printf("%zu ", offsetof(mpm_config, {f}));')
    src.append('printf("\\n");')
    for f in fields_pts:
        src.append(f'printf("%zu ", offsetof(mpm_points, {f}));')
    src.append('printf("\\n"); return 0; }')
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    s_cfg, s_pts = map(int, lines[0].split())
    assert s_cfg == C.sizeof(A.MpmConfig)
    assert s_pts == C.sizeof(A.MpmPoints)
    assert list(map(int, lines[1].split())) == [getattr(A.MpmConfig, f).offset for f in fields_cfg]
    assert list(map(int, lines[2].split())) == [getattr(A.MpmPoints, f).offset for f in fields_pts]


def test_nodes_layout_matches_c(tmp_path):
    """sizeof / offsetof of mpm_nodes from a C compile of the header."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    fields = [f[0] for f in A.MpmNodes._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpm_auxetics.h"', "int main(void){",
           'printf("%zu\\n", sizeof(mpm_nodes));']
    src += [f'printf("%zu ", offsetof(mpm_nodes, {f}));' for f in fields]
    src.append('printf("\\n"); return 0; }')
    c = tmp_path / "nodes.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "nodes"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert int(lines[0]) == C.sizeof(A.MpmNodes)
    assert list(map(int, lines[1].split())) == [getattr(A.MpmNodes, f).offset for f in fields]


def test_grid_nodes_argument_errors_without_device():
    lib = A.lib()
    nd = A.MpmNodes()
    assert lib.mpm_grid_nodes(None, C.byref(nd)) == A.MPM_E_ARG


def test_constitutive_batch_layout_matches_c(tmp_path):
    """sizeof / offsetof of mpm_cr_batch (the ctypes field ``yield_`` is the
    C field ``yield``) from a C compile of the header."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    fields = [f[0] for f in A.MpmCrBatch._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpm_auxetics.h"', "int main(void){"]
    src.append('printf("%zu\\n", sizeof(mpm_cr_batch));')
    for f in fields:
        src.append(f'printf("%zu ", offsetof(mpm_cr_batch, {f.rstrip("_")}));')
    src.append('printf("\\n"); return 0; }')
    c = tmp_path / "crlayout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "crlayout"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert int(lines[0]) == C.sizeof(A.MpmCrBatch)
    assert list(map(int, lines[1].split())) == [getattr(A.MpmCrBatch, f).offset for f in fields]


def test_constitutive_argument_errors_without_device():
    """Laws the reference never dispatches, unknown laws and missing arrays
    are argument errors, reported before any device is needed."""
    lib = A.lib()
    b = A.MpmCrBatch()
    b.n = 1
    for law in (A.MPM_LAW_VISCOELASTIC, A.MPM_LAW_IDEALGAS):
        assert lib.mpm_constitutive(law, C.byref(b), 0) == A.MPM_E_ARG
        assert b"not dispatched" in lib.mpm_last_error(None)
    assert lib.mpm_constitutive(7, C.byref(b), 0) == A.MPM_E_ARG
    assert b"unknown" in lib.mpm_last_error(None)
    assert lib.mpm_constitutive(A.MPM_LAW_ELASTIC, C.byref(b), 0) == A.MPM_E_ARG   # no arrays
    assert lib.mpm_constitutive(A.MPM_LAW_ELASTIC, None, 0) == A.MPM_E_ARG
    b.n = 0
    assert lib.mpm_constitutive(A.MPM_LAW_ELASTIC, C.byref(b), 0) == A.MPM_OK


def _config(**kw):
    cfg = A.MpmConfig()
    for a in range(3):
        cfg.length[a] = 0.01
        cfg.cells[a] = 10
    cfg.dt_max = 1e-7
    cfg.time_end = 1.0
    cfg.precision = A.MPM_FP32
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def test_create_rejects_bad_arguments_before_touching_a_device():
    lib = A.lib()
    h = C.c_void_p()
    assert lib.mpm_create(None, C.byref(h)) == A.MPM_E_ARG
    assert lib.mpm_create(C.byref(_config(dt_max=0.0)), C.byref(h)) == A.MPM_E_ARG
    assert b"dt_max" in lib.mpm_last_error(None)
    assert lib.mpm_create(C.byref(_config(precision=7)), C.byref(h)) == A.MPM_E_ARG
    bad = _config()
    bad.cells[1] = 0
    assert lib.mpm_create(C.byref(bad), C.byref(h)) == A.MPM_E_ARG
    assert lib.mpm_create(C.byref(_config(n_timeline=2)), C.byref(h)) == A.MPM_E_ARG
    # cell indices enter 2^20-bounded wave reductions of the contact step
    huge = _config()
    huge.cells[0] = (1 << 20) - 1
    assert lib.mpm_create(C.byref(huge), C.byref(h)) == A.MPM_E_ARG
    assert b"2^20" in lib.mpm_last_error(None)
    assert not h.value


def test_create_rejects_fp32_bricks_beyond_the_fixed_point_range():
    """FP32 state keeps positions as Q6.26 cells from the brick origin (int32:
    |u| < 32 cells).  A brick edge above 28 cells would let a drifting point
    overflow it, so mpm_create refuses such bricks in FP32 (before any device
    call) and accepts them in FP64, which stores absolute positions."""
    lib = A.lib()
    h = C.c_void_p()
    big = _config(precision=A.MPM_FP32)
    big.cells[2] = 64
    big.brick[0], big.brick[1], big.brick[2] = 3, 3, 32
    assert lib.mpm_create(C.byref(big), C.byref(h)) == A.MPM_E_ARG
    assert b"at most 28 cells" in lib.mpm_last_error(None)
    assert not h.value
    for prec, edge in ((A.MPM_FP32, 28), (A.MPM_FP64, 32)):
        cfg = _config(precision=prec)
        cfg.cells[2] = 64
        cfg.brick[0], cfg.brick[1], cfg.brick[2] = 3, 3, edge
        rc = lib.mpm_create(C.byref(cfg), C.byref(h))
        assert rc in (A.MPM_OK, A.MPM_E_NO_DEVICE), lib.mpm_last_error(None)
        if rc == A.MPM_OK:
            lib.mpm_destroy(h)


def test_create_without_a_device_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = A.lib()
    h = C.c_void_p()
    assert lib.mpm_create(C.byref(_config()), C.byref(h)) == A.MPM_E_NO_DEVICE
    assert lib.mpm_last_error(None) == b"no HIP device"
    assert not h.value


def test_null_handle_calls_are_argument_errors():
    lib = A.lib()
    fin = C.c_int32()
    assert lib.mpm_step(None, C.c_double(1e-6), 0, C.byref(fin)) == A.MPM_E_ARG
    out = (C.c_double * 12)()
    assert lib.mpm_checks(None, out) == A.MPM_E_ARG
    lib.mpm_destroy(None)   # no-op


def _build_host_demo(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "host_engine_demo"
    libdir = os.path.dirname(A.LIB_PATH)
    subprocess.run([gxx, "-std=c++17", "-O1", os.path.join(ROOT, "tests", "host_engine_demo.cpp"), "-o", str(exe),
                    "-L", libdir, "-lmpm_auxetics", f"-Wl,-rpath,{libdir}"], check=True)
    return exe


def test_cpp_host_mirror_builds_and_reports_no_device(tmp_path):
    """The C++ PhysicsEngine mirror compiles against the header, links the
    library and surfaces MPM_E_NO_DEVICE as an exception without a GPU."""
    import torch
    exe = _build_host_demo(tmp_path)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible (the GPU test runs the demo)")
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "no HIP device" in r.stdout


def test_constitutive_relation_keeps_the_undispatched_signatures():
    """ConstitutiveRelation.h:16-34: the ViscoElastic and IdealGas methods
    exist with the reference's arguments and raise, as the step never
    dispatches them (…_MP.cpp:332-339)."""
    import inspect

    from code_mpm_auxetics_amd.constitutive import ConstitutiveRelation
    c = ConstitutiveRelation()
    assert list(inspect.signature(c.calculateIncrement_ViscoElastic_6D).parameters) == \
        ["dE", "dNu", "dEta", "d6Stress_Current", "d6Strain_New", "d6Strain_Rate_New"]
    assert list(inspect.signature(c.calculateIncrement_IdealGass).parameters) == \
        ["dHeatCapacityRatio", "dSpecificHeat", "dShearViscosity", "dBulkViscosity", "d6Stress_Current",
         "dDensity", "dTemperature", "d6Strain_Rate"]
    with pytest.raises(NotImplementedError):
        c.calculateIncrement_ViscoElastic_6D(1.0, 0.3, 1.0, [0] * 6, [0] * 6, [0] * 6)
    with pytest.raises(NotImplementedError):
        c.calculateIncrement_IdealGass(1.4, 1.0, 0.0, 0.0, [0] * 6, 1.0, 300.0, [0] * 6)
    assert c.d6StressIncrement.shape == (6,) and c.d_J2 == 0.0


def test_cpp_constitutive_mirror_compiles(tmp_path):
    """The C++ ConstitutiveRelation mirror (ConstitutiveRelation.h:16-34
    method names and output members) compiles and links against the
    library; no compute call is made."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "cr.cpp"
    src.write_text('#include "code_mpm_auxetics_amd/host/ConstitutiveRelation.h"\n'
                   "int main(int argc, char **) {\n"
                   "    mpm_host::ConstitutiveRelation cr;\n"
                   "    if (argc > 1) { const double s[6] = {1, 0, 0, 0, 0, 0}; cr.calculateState_J2(s); }\n"
                   "    return cr.d6StressIncrement[0] == 0.0 ? 0 : 1;\n"
                   "}\n")
    exe = tmp_path / "cr"
    libdir = os.path.dirname(A.LIB_PATH)
    subprocess.run([gxx, "-std=c++17", "-I", ROOT, str(src), "-o", str(exe), "-L", libdir, "-lmpm_auxetics",
                    f"-Wl,-rpath,{libdir}"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_async_step_argument_errors_without_device():
    lib = A.lib()
    done = C.c_int32()
    assert lib.mpm_step_async(None, C.c_double(1e-6), 0) == A.MPM_E_ARG
    assert lib.mpm_step_poll(None, C.byref(done)) == A.MPM_E_ARG
    assert lib.mpm_step_wait(None, C.byref(done)) == A.MPM_E_ARG
