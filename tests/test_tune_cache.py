"""The autotune cache file (lbm_amd/csrc/tune_cache.h), host code only: what
lbm_autotune_cached and lbm_batch_autotune_cached read and write.  The file
must stay valid JSON through every store, keep other keys, read back what it
stored (a candidate that could not run as null), and ignore what it did not
write.  No reference counterpart: the reference has no autotuner."""
import json
import os
import subprocess

import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    exe = tmp_path_factory.mktemp("tc") / "tune_cache_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    os.path.join(ROOT, "tests", "cpp", "tune_cache_check.cpp"), "-o", str(exe)],
                   check=True, timeout=120)

    def run(*args):
        r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=60)
        return r.returncode, r.stdout.strip()
    return run


KEY_A = "gfx950:sramecc+:xnack-/256cu/grid/512x256/f32/ref/v10"
KEY_B = "gfx950:sramecc+:xnack-/256cu/batch/512d/256Mcells/f16e/ref/v6"


def test_store_and_lookup_round_trip(tool, tmp_path):
    path = tmp_path / "tune.json"
    assert tool("lookup", path, KEY_A) == (0, "miss")  # no file yet
    assert tool("store", path, KEY_A, 8, 20, 0.5, 0.25, "inf")[0] == 0
    assert tool("lookup", path, KEY_A) == (0, "8 20 0.5 0.25 inf")
    assert tool("store", path, KEY_B, 3, 10, 1.5)[0] == 0
    data = json.loads(path.read_text())  # valid JSON, both keys kept
    assert data[KEY_A] == {"variant": 8, "steps": 20, "ms": [0.5, 0.25, None]}
    assert data[KEY_B] == {"variant": 3, "steps": 10, "ms": [1.5]}
    # replacing a key keeps the others
    assert tool("store", path, KEY_A, 9, 30, 0.125)[0] == 0
    data = json.loads(path.read_text())
    assert data[KEY_A]["variant"] == 9 and data[KEY_B]["variant"] == 3
    assert tool("lookup", path, KEY_B) == (0, "3 10 1.5")
    assert not [p for p in os.listdir(tmp_path) if ".tmp." in p]  # temporary file renamed away


def test_foreign_and_damaged_lines_are_ignored(tool, tmp_path):
    path = tmp_path / "tune.json"
    path.write_text("\n".join([
        "{",
        '"gfx950/256cu/grid/64x64/f32/ref/v10": {"variant": 2, "steps": 5, "ms": [0.1, 0.2]},',
        '"bad key with spaces": {"variant": 1, "steps": 5, "ms": []},',
        '"gfx950/256cu/grid/128x64/f32/ref/v10": {"variant": -1, "steps": 5, "ms": []},',
        '"gfx950/256cu/grid/256x64/f32/ref/v10": {"variant": 4, "steps": 5, "ms": [0.1, oops]},',
        '"gfx950/256cu/grid/512x64/f32/ref/v10": {"variant": 4, "steps": 5',
        "garbage",
        "}",
    ]))
    assert tool("lookup", path, "gfx950/256cu/grid/64x64/f32/ref/v10") == (0, "2 5 0.1 0.2")
    for k in ("gfx950/256cu/grid/128x64/f32/ref/v10", "gfx950/256cu/grid/256x64/f32/ref/v10",
              "gfx950/256cu/grid/512x64/f32/ref/v10"):
        assert tool("lookup", path, k) == (0, "miss"), k
    assert tool("store", path, KEY_A, 0, 1, 1.0)[0] == 0
    data = json.loads(path.read_text())  # rewritten: only the entries that read back
    assert sorted(data) == sorted([KEY_A, "gfx950/256cu/grid/64x64/f32/ref/v10"])


def test_unwritable_path_and_bad_key_fail(tool, tmp_path):
    assert tool("store", tmp_path / "no" / "such" / "dir.json", KEY_A, 1, 1, 1.0)[0] == 1
    assert tool("store", tmp_path / "t.json", 'a"b', 1, 1, 1.0)[0] == 1
    assert not (tmp_path / "t.json").exists()


def test_shape_class_and_key_cleaning(tool):
    assert [tool("pow2", n)[1] for n in (1, 2, 3, 160, 400, 4096, 4097)] == \
        ["1", "2", "4", "256", "512", "4096", "8192"]
    assert tool("clean", 'gfx950:sramecc+:xnack-/x "y"')[1] == "gfx950:sramecc+:xnack-/x__y_"
