"""bench.py's command-line contract on the CPU: one GPU only, and the LBVH
byte model it reports against."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_build_byte_model():
    sys.path.insert(0, str(ROOT))
    import bench

    # the build's contract model: bounds 36 + Morton 40 + sort 64 + Karras 20 + refit 164
    assert bench.BUILD_BYTES_PER_TRI == 36 + 40 + 64 + 20 + 164
    stages = {name: (kernel, bpt, launches) for name, kernel, bpt, launches in bench.STAGES}
    assert stages["hier_local"] == ("k_hier_local", 184.0, 1)
    assert stages["sort"][1:] == (64.0, 4)
    # k_bounds: soup 36 in, centre 12 out (the hierarchy pass gathers the triangle for the leaf box);
    # k_morton: centre 12 in, key 4 out (the hierarchy pass writes the arrival words of the splits it
    # leaves open)
    assert stages["bounds"][1] == 36 + 12
    assert stages["morton"][1] == 12 + 4


def test_build_byte_model_follows_the_key_width():
    sys.path.insert(0, str(ROOT))
    import bench

    # 10 bits per axis: the model above, four 8-bit passes over 32-bit keys
    stages, whole = bench.lbvh_model(10)
    assert stages == bench.STAGES and whole == bench.BUILD_BYTES_PER_TRI
    # 20 bits per axis (the reference's default): 60-bit keys in 64-bit words,
    # eight passes of 8-B keys and 4-B payloads read and written, and 4 B
    # more in the Morton pass's key write and the hierarchy's key reads
    stages, whole = bench.lbvh_model(20)
    s = {name: (bpt, launches) for name, _, bpt, launches in stages}
    assert s["sort"] == (8 * 2 * (8 + 4), 8)
    assert s["morton"][0] == 12 + 8 and s["hier_local"][0] == 184 + 4
    assert whole == bench.BUILD_BYTES_PER_TRI - 64 + 8 * 24 + 8
    # 5 bits per axis: two passes of 32-bit keys
    stages, whole = bench.lbvh_model(5)
    assert {name: launches for name, _, _, launches in stages}["sort"] == 2
    assert whole == bench.BUILD_BYTES_PER_TRI - 64 + 2 * 16
