"""The host runtime (equation parser, expression evaluator, code generator, Sobol tables, argument
checks of the C ABI) built with AddressSanitizer and UndefinedBehaviorSanitizer, running
tests/test_host.py: any out-of-bounds access, use after free or undefined behaviour in the host
code fails the run. CPU only; the device code is not instrumented."""
import os
import shutil
import subprocess
import sys

from conftest import ROOT


def test_host_suite_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    sys.path.insert(0, os.path.join(ROOT, "sde-sim-rs_amd"))
    import build as pkg_build  # sde-sim-rs_amd/build.py
    pkg = tmp_path / "sde_sim_rs"
    shutil.copytree(os.path.join(ROOT, "sde-sim-rs_amd", "sde_sim_rs"), pkg,
                    ignore=shutil.ignore_patterns("*.so", "__pycache__"))
    pkg_build.build_sanitized(str(pkg / "libsdesim.so"))
    syms = subprocess.run(["nm", "-D", "--undefined-only", str(pkg / "libsdesim.so")], stdout=subprocess.PIPE,
                          text=True, check=True).stdout
    assert "__asan_report_load8" in syms and "__ubsan_handle_" in syms  # instrumented host code
    env = pkg_build.sanitizer_env()
    env["SDESIM_TEST_PACKAGE_ROOT"] = str(tmp_path)
    env["SDESIM_CACHE_DIR"] = str(tmp_path / "kernels")
    # The hipRTC compile cases exercise the compiler library, not this code; the rest of the
    # suite runs unchanged.
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_host.py"), "-k", "not compile and not bench_refuses"],
                       env=env, cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-6000:]
    print(r.stdout.strip().splitlines()[-1])
    assert "sde_sim_rs/libsdesim.so" not in r.stdout
    # The run loaded the sanitized copy, not the in-tree library.
    chk = subprocess.run([sys.executable, "-c", "import sde_sim_rs._native as n; print(n.LIB_PATH)"], env=env,
                         cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert chk.stdout.strip().endswith(str(pkg / "libsdesim.so")), chk.stdout[-2000:]
