"""bench.py's contract on the CPU: it runs on one GPU only, and the HBM
traffic it reports per launch comes from the committed counter summary,
which must parse and stay close to the face kernel's algorithmic bytes."""

import importlib.util
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_traffic_summary_parses():
    bench = _bench()
    for kernel in ("element", "face"):
        for key in ("f64", "f32"):
            v = bench.measured_traffic(kernel, key)
            assert v is not None and v > 0, (kernel, key)
    # FP64 moves twice the FP32 bytes, to within the counters' granularity
    for kernel in ("element", "face"):
        ratio = bench.measured_traffic(kernel, "f64") / bench.measured_traffic(kernel, "f32")
        assert 1.8 < ratio < 2.1, (kernel, ratio)


def test_face_traffic_is_the_algorithmic_traffic():
    # TGV 32^3 periodic hexahedra, p=3: 3 * 32^3 interior faces, 64 nodes
    # and 16 face-support nodes per side, 5 variables
    bench = _bench()
    faces = 3 * 32 ** 3
    for key, itemsize in (("f64", 8), ("f32", 4)):
        algo = bench.face_bytes(64, 16, 5, itemsize) * faces
        measured = bench.measured_traffic("face", key)
        assert abs(measured / algo - 1.0) < 0.1, (key, measured, algo)
