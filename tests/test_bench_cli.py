"""bench.py on the CPU: the batch workload it times is the documented one
(64 seeded Au/Ag spheres and ellipsoids at 16 wavelengths, 1,024 solves,
3,658,944 dipoles, N from 17 to 13,949), and a request for more than one GPU
ends with a message and exit status 2 before any device work."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_batch_workload_is_the_documented_one():
    import bench
    probs = bench.batch_problems()
    assert len(probs) == 1024
    n = np.array([p["n"] for p in probs])
    assert n.sum() == 3658944 and n.min() == 17 and n.max() == 13949
    # 16 wavelengths per particle, 400..800 nm in water: k = 2 pi 1.33 / lambda
    lam = 400.0 + 400.0 * np.arange(16) / 15
    k = np.array([p["k"] for p in probs[:16]])
    assert np.allclose(k, 2 * np.pi * 1.33 / lam, rtol=1e-14)
    # consecutive solves of one particle share its positions
    assert all(probs[q]["positions"] is probs[0]["positions"] for q in range(16))
    assert all(p["use_fcd"] and p["cell_size"] == 2.0 for p in probs)
