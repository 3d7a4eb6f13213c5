"""Runs the C++ tests of the host API (tests/cpp, built by
__graft_entry__.build() / `make -C tests/cpp`)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "cpp", "build")


def _binary(name, build_if_missing):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        if not build_if_missing:
            pytest.fail(f"{path} is not built (run __graft_entry__.build())")
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), f"build/{name}"],
                       check=True)
    return path


def _run(path, cwd):
    p = subprocess.run([path], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       cwd=cwd, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout[-3000:]


def test_host_api_cpu(orc, tmp_path):
    _run(_binary("host_cpu_test", True), tmp_path)


@pytest.mark.gpu
def test_host_api_gpu(tmp_path):
    _run(_binary("host_gpu_test", False), tmp_path)


def test_unknown_advection_name_aborts(tmp_path):
    """advect_factory.cpp:44-47: an unknown name aborts listing the names."""
    src = tmp_path / "bad_name.cpp"
    src.write_text('#include "uvlm_amd/host/advect.h"\n'
                   'int main() { UVLM::advect::AdvectFactory("rk4"); return 0; }\n')
    root = os.path.dirname(HERE)
    lib = os.path.join(root, "uvlm_amd", "lib")
    exe = tmp_path / "bad_name"
    subprocess.run(["g++", "-std=c++17", f"-I{root}", str(src), "-o", str(exe), f"-L{lib}",
                    "-luvlm", "-luvlm_gpu", f"-Wl,-rpath,{lib}"], check=True)
    p = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode != 0
    assert "Invalid advection name! Please choose from following: [AB2, RK2, RK4, RK4classic, euler]" in p.stdout


@pytest.mark.parametrize("cols,wings,msg", [(7, 1, "columns must be even"),
                                            (8, 0, "num_wings out of range")])
def test_legacy_run_rejects_bad_specs(cols, wings, msg):
    """uvlm_host_legacy_run checks the case before any device work (the
    reference aborts on odd columns, simple_simulator.cpp:136)."""
    from uvlm_amd import workloads as W
    from uvlm_amd import _native as N
    from uvlm_amd.gpu import UVLMError, case_spec, legacy_run
    case = W.plate_c1(steps=2, rows=2, cols=cols)
    if wings == 1:
        with pytest.raises(UVLMError, match=msg):
            legacy_run(case, 2)
    else:
        import ctypes as C
        spec = case_spec(case)
        spec.num_wings = 0
        L = N.host_lib()
        nw = C.c_size_t(0)
        assert L.uvlm_host_legacy_run(C.byref(spec), 2, None, None, None, None,
                                      C.byref(nw)) != N.OK
        assert msg in L.uvlm_host_last_error().decode()
