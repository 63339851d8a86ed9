"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

  * the CPU restatement (oracle/gft_oracle.cpp): every entry point over every
    model option on small lattices (oracle/asan_check.cpp, make -C oracle asan);
  * the C++ host API (gftevolver_amd/host): gft_run's configuration reader, the
    boundary-region construction and the RandomUniform initial state, through
    `gft_run ... dump_config` on the configurations of the reference program's
    golden runs, plus an invalid configuration (the error path).  The GPU library
    is linked but not called by dump_config.

Both are built with g++ (host code only) and the sanitizer runtimes linked
statically; a report aborts the process, so a non-zero status fails the test."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import pytest

from gftevolver_amd._abi import TWO_HDM_POTENTIAL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
TEMPLATE = os.path.join(ORACLE, "_ref", "cfg")
LIBGFT = os.path.join(ROOT, "gftevolver_amd", "libgft.so")
SAN = ["-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-static-libasan", "-static-libubsan"]


def _has_asan():
    if not shutil.which("g++"):
        return False
    r = subprocess.run(["g++", "-fsanitize=address", "-static-libasan", "-x", "c++", "-", "-o", os.devnull],
                       input="int main(){return 0;}", capture_output=True, text=True)
    return r.returncode == 0


needs_asan = pytest.mark.skipif(not _has_asan(), reason="g++ with libasan not available")


@needs_asan
def test_oracle_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True, capture_output=True, timeout=600)
    r = subprocess.run([os.path.join(ORACLE, "_asan", "asan_check")], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0 and "cases clean" in r.stdout


@needs_asan
@pytest.mark.skipif(not (os.path.isdir(TEMPLATE) and os.path.exists(LIBGFT)),
                    reason="reference config template or libgft.so not built")
def test_host_config_path_clean_under_asan_ubsan(tmp_path):
    sys.path.insert(0, ORACLE)
    import refcfg
    exe = str(tmp_path / "gft_run_asan")
    host = os.path.join(ROOT, "gftevolver_amd", "host")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17"] + SAN +
                   ["-o", exe, os.path.join(host, "gft_host.cpp"), os.path.join(host, "gft_run.cpp"),
                    "-I" + host, "-I" + os.path.join(ROOT, "include"), "-L" + os.path.dirname(LIBGFT), "-lgft",
                    "-Wl,-rpath," + os.path.dirname(LIBGFT)], check=True, capture_output=True, timeout=600)
    cases = {
        "sm": dict(model=0, ns=4),
        "2hdm": dict(model=1, ns=8, pot2=TWO_HDM_POTENTIAL),
        "sm_gen": dict(model=0, ns=4, gen=True, nvd=4),
        "sm_fixed_xz": dict(model=0, ns=4, faces=(1, 1, 3, 3, 1, 1)),
        "sm_fixed_y": dict(model=0, ns=4, faces=(3, 3, 1, 1, 3, 3)),
        "2hdm_g5": dict(model=1, ns=8, pot2=TWO_HDM_POTENTIAL, nvd=6),
        "2hdm_g7_gen": dict(model=1, ns=8, pot2=TWO_HDM_POTENTIAL, gen=True, nvd=7),
        "global_n3": dict(model=2, ns=3, nvd=1, gauss=False),
        "global_n8": dict(model=2, ns=8, pot2=TWO_HDM_POTENTIAL, nvd=1, gauss=False),
        "dumps_on": dict(model=0, ns=4, output_fields=True),
    }
    for name, c in cases.items():
        td = str(tmp_path / name)
        refcfg.prepare(td)
        refcfg.write_model(td, c["model"], 0.65, 0.35, pot2=c.get("pot2"), gen=c.get("gen", False))
        refcfg.write_lattice(td, (6, 5, 7), (0.5, 0.5, 0.5), 12, 0.05, c["ns"], ic_seed=7, ic_range=(-0.8, 0.8),
                             efreq=3, gfreq=2, faces=c.get("faces", (3, 3, 3, 3, 3, 3)), nvd=c.get("nvd", 5),
                             gauss=c.get("gauss", True), output_fields=c.get("output_fields", False))
        r = subprocess.run([exe, os.path.join(td, "cfg"), os.path.join(td, "data"), "dump_config"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (name, r.stderr[-3000:])
        assert r.stdout.startswith("n 6 5 7"), name
    # The error path: an invalid potential type throws the reference's message.
    td = str(tmp_path / "bad")
    refcfg.prepare(td)
    refcfg.write_model(td, 0, 0.65, 0.35)
    refcfg.write_lattice(td, (6, 5, 7), (0.5, 0.5, 0.5), 12, 0.05, 4)
    refcfg.set_values(td, "Core", "Model.cfg", [7, 2, 1, 0.0, 0, 0, 0.0])
    r = subprocess.run([exe, os.path.join(td, "cfg"), os.path.join(td, "data"), "dump_config"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "The chosen potential type (7) is invalid." in r.stderr
    assert "Sanitizer" not in r.stderr
