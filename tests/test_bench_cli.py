"""The benchmark's command-line contract on the CPU: one GPU only, and the
frozen work counts it prices the roofline with."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_flops_header_counts():
    fl = bench.flops_header()
    # pair-loop counts: dX 3, r^2 5, sqrt+div 2, rho 1, kernel 8 / 12, dX x Gamma 9,
    # -1/(4 pi r^3) 3, g a 1, U 6, radial aux 9, J radial 18, J Kronecker 12
    common = 3 + 5 + 2 + 1 + 9 + 3 + 1 + 6 + 9 + 18 + 12
    assert fl["FLOPS_UJ_PAIR_GAUSSIANERF"] == common + 8 == 77
    assert fl["FLOPS_UJ_PAIR_WINCKELMANS"] == common + 12 == 81
    # fused RK substep: 30 doubles read + 14 written
    assert fl["BYTES_RK_SUBSTEP_PER_PARTICLE"] == (30 + 14) * 8
    assert fl["PEAK_FP32_VECTOR_TFLOPS"] == 157.3 and fl["PEAK_HBM_GBS"] == 8000.0
