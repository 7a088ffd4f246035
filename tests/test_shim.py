"""The drop-in shim on the CPU: fargo_thorin_amd/shim/algogas_shim.c
compiles against the reference's own headers (src_main/fargo.h), links with
the reference's objects in place of SourceEuler.c's AlgoGas (made weak), and
the binary it makes refuses to run without a GPU rather than stepping on the
CPU."""
import os
import shutil
import subprocess

import pytest

from tests.golden.make_golden import CASES, prepare_case, run_args
from tests.helpers import short_tmpdir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src_main"
LIB = os.path.join(ROOT, "fargo_thorin_amd", "libthorin_amd.so")
BIN = os.path.join(ROOT, "oracle", "_ref", "thorin_gpu")


@pytest.fixture
def run_dir():
    with short_tmpdir() as d:
        yield d


def _have_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="module")
def shim_binary():
    if not os.path.isdir(REF_SRC):
        pytest.skip("the reference's sources are not here")
    if not os.path.exists(LIB):
        pytest.skip("libthorin_amd.so is not built")
    r = subprocess.run(["make", "-f", "oracle/ref.mk", "oracle/_ref/thorin_gpu"], cwd=ROOT,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return BIN


def _nm(path, *flags):
    return subprocess.run(["nm", *flags, path], capture_output=True, text=True, check=True).stdout


def _symbols(text, name):
    return [ln.split()[-2] for ln in text.splitlines() if ln.split() and ln.split()[-1] == name]


def test_reference_algogas_is_weak_and_the_shims_is_linked(shim_binary):
    weak = os.path.join(ROOT, "oracle", "_ref", "main_SourceEuler_weak.o")
    assert _symbols(_nm(weak), "AlgoGas") == ["W"]
    shim = os.path.join(ROOT, "oracle", "_ref", "shim_algogas.o")
    s = _nm(shim)
    assert _symbols(s, "AlgoGas") == ["T"]
    for f in ("thorin_ctx_create", "thorin_algogas", "thorin_init_state", "thorin_init_pebbles",
              "thorin_download_field", "thorin_upload_field", "thorin_get_state", "thorin_set_state"):
        assert _symbols(s, f) == ["U"], f
    # the reference's own N-body and output routines, called from the shim
    for f in ("AdvanceSystemRebound", "RotatePsys", "SynchronizeFargoRebound", "WriteDiskPolar"):
        assert _symbols(s, f) == ["U"], f
    b = _nm(shim_binary)
    assert _symbols(b, "AlgoGas") == ["T"]
    # the strong definition is the shim's: its helpers sit beside it
    assert _symbols(b, "shim_nbody") == ["t"]


def test_shim_binary_loads_the_in_tree_library(shim_binary):
    r = subprocess.run(["ldd", shim_binary], capture_output=True, text=True, check=True)
    line = [ln for ln in r.stdout.splitlines() if "libthorin_amd.so" in ln]
    assert line, r.stdout
    assert os.path.realpath(line[0].split("=>")[1].split("(")[0].strip()) == os.path.realpath(LIB)


@pytest.mark.skipif(_have_gpu(), reason="with a GPU the binary steps (the GPU tests run it)")
def test_shim_binary_refuses_to_run_without_a_gpu(shim_binary, run_dir):
    name = "crash_cold_48x96"
    cfg, planets, _ = CASES[name]
    par, outdir, _, _ = prepare_case(name, cfg, planets, run_dir, binary=shim_binary)
    r = subprocess.run([shim_binary] + run_args(cfg, par), cwd=run_dir, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1
    assert "GPU hydro step: context: no HIP device" in r.stdout
    # the reference's set-up wrote output 0; no DT was taken
    assert os.path.exists(os.path.join(outdir, "gasdens0.dat"))
    assert not os.path.exists(os.path.join(outdir, "gasdens1.dat"))
    assert shutil.which("nm")
