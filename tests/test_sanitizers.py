"""Host code under AddressSanitizer and UBSan (CPU only).

The C-ABI host side (topology, metadata, batching, argument checks) and the
CPU oracle are rebuilt with -fsanitize=address,undefined (device code is
unchanged) and exercised in a child interpreter with the sanitizer runtime
preloaded: refinement regularisation and maps, scene level policy, the
prepare path up to the device upload (which fails without a GPU after the
host metadata is built), argument errors, and a small oracle solve.  Any
sanitizer report fails the test."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT, has_gpu

CHILD = r"""
import os, sys
sys.path.insert(0, os.path.join(sys.argv[1], "msbg-flip_amd"))
sys.path.insert(0, os.path.join(sys.argv[1], "tests"))
import numpy as np
import msbgx
from oracle_lib import build_oracle

for cfg in (msbgx.config_c4(n=64), msbgx.config_c2(n=128, bsx=32), msbgx.config_c5(n=128)):
    g = msbgx.build_grid(cfg)
    try:
        g.prepare_pressure_solve_flip(scene=cfg.scene)
    except msbgx.MsbgxError as e:
        assert e.rc == -3, e          # host metadata built, device upload refused
    for mg in range(cfg.levels):
        g.block_info(mg)
        g.relax_list(mg)
    g.destroy()
n = 32
g = msbgx.MultiresSparseGrid.create(n, n, n, 16, 2)
flags = np.zeros((n, n, n), np.uint16)
w = [np.ones((n, n, n), np.float32) for _ in range(3)]
rhs = np.zeros((n, n, n), np.float32)
rhs[3, 3, 3] = np.nan
try:
    g.prepare_pressure_solve_flip(dense=(flags, *w, rhs))
except msbgx.MsbgxError as e:
    assert e.rc == -1
for bad in ((100, 64, 64, 16, 3), (64, 64, 64, 12, 2), (64, 64, 64, 16, 9)):
    try:
        msbgx.MultiresSparseGrid.create(*bad)
    except msbgx.MsbgxError:
        pass
cfg = msbgx.config_c4(n=64)
o = build_oracle(cfg, threads=2)
o.prepare_scene(cfg.scene)
it, rr, _ = o.solve(max_iter=40, tol=1e-6)
assert it > 0 and rr < 1e-3
print("sanitized run ok")
"""


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.skipif(has_gpu(), reason="the device-free prepare path is exercised without a GPU")
def test_host_code_is_clean_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    ubsan = _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("sanitizer runtimes not available")
    subprocess.run(["make", "asan"], cwd=os.path.join(ROOT, "msbg-flip_amd"), check=True, capture_output=True)
    subprocess.run(["make", "asan"], cwd=os.path.join(ROOT, "oracle"), check=True, capture_output=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan}:{ubsan}",
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1",
        "MSBGX_LIB": os.path.join(ROOT, "msbg-flip_amd", "_build", "asan", "libmsbgx_asan.so"),
        "MSBG_ORACLE_SO": os.path.join(ROOT, "oracle", "_build", "libmsbg_oracle_asan.so"),
        "OMP_NUM_THREADS": "2",
    })
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=600)
    report = r.stdout[-3000:] + r.stderr[-6000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, report
    assert r.returncode == 0 and "sanitized run ok" in r.stdout, report
