"""Concurrency of the candidate labeling (k_cand_local, k_cand_merge, k_cand_flatten in
csrc/ccl.hip), checked on the CPU with a host-threaded model of the three kernels
(tests/ccl_model/ccl_model.cpp): work-groups as teams of threads with barriers, several
at once, random yields between memory operations, random tile shapes, record orders,
volume sizes (widths and heights of 1 and 2 included) and leftover contents of the
parent array.  Every run's labels, root bits and region-spanning bits are compared with
the oracle's sequential union-find.

- atomic mode (the kernels' accesses: relaxed atomic loads and stores of the parent array
  beside the atomic minimum) gives the oracle's labels, and is clean under
  ThreadSanitizer;
- plain mode (the earlier kernels: plain loads and stores beside the atomic minimum) is a
  data race, which ThreadSanitizer reports;
- stale mode serves the plain accesses from a private write-back cache per lane (stale
  loads, stores that reach memory late and in any order) and still gives the oracle's
  labels, so the union-find does not depend on the accesses being coherent;
- each of five planted mistakes (a face entry not written, outside neighbours counted as
  known, path halving in the flatten, a missing carry of a wrapped row, a lost race of
  the minimum not retried) makes the model fail, so the comparison can fail."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "ccl_model", "ccl_model.cpp")
ORACLE = os.path.join(ROOT, "oracle", "hdog_oracle.cpp")


def build(tmp, name, *flags):
    exe = os.path.join(tmp, name)
    subprocess.run(["g++", "-std=c++17", "-pthread", "-ffp-contract=off", *flags, "-o", exe,
                    SRC, ORACLE], check=True)
    return exe


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    return str(tmp_path_factory.mktemp("ccl_model"))


@pytest.fixture(scope="module")
def model(tmp):
    return build(tmp, "ccl_model", "-O2")


def run(exe, mode, iters, seed, teams=4, lanes=4, timeout=300):
    return subprocess.run([exe, mode, str(iters), str(seed), str(teams), str(lanes)],
                          capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("mode,seed,teams,lanes", [("atomic", 1, 4, 4), ("atomic", 2, 2, 8),
                                                   ("atomic", 3, 8, 2), ("stale", 4, 4, 4),
                                                   ("stale", 5, 8, 3), ("plain", 6, 4, 4)])
def test_labels_match_the_oracle_under_random_interleavings(model, mode, seed, teams, lanes):
    r = run(model, mode, 1500, seed, teams, lanes)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    _, vols, cand, unions = r.stdout.split()
    assert int(vols) == 1500 and int(cand) > 100000 and int(unions) > 10000


@pytest.fixture(scope="module")
def model_tsan(tmp):
    return build(tmp, "ccl_model_tsan", "-O1", "-g", "-fsanitize=thread")


def test_atomic_accesses_are_race_free_under_tsan(model_tsan):
    r = run(model_tsan, "atomic", 150, 7, 3, 3)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[:4000]
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr[:2000]


def test_plain_accesses_are_a_data_race_under_tsan(model_tsan):
    # the parent-array accesses the merge made before they became atomics
    r = run(model_tsan, "plain", 40, 8, 3, 3)
    assert "WARNING: ThreadSanitizer: data race" in r.stderr
    assert "find_root" in r.stderr or "kernel_merge" in r.stderr


@pytest.mark.parametrize("plant,mode", [(1, "atomic"), (2, "atomic"), (3, "stale"), (4, "atomic"),
                                        (5, "atomic")])
def test_planted_mistakes_are_caught(tmp, plant, mode):
    exe = build(tmp, f"ccl_model_plant{plant}", "-O2", f"-DPLANT={plant}")
    r = run(exe, mode, 6000, 11 + plant, 4, 4)
    assert r.returncode != 0, f"planted mistake {plant} went unnoticed: {r.stdout}"
    assert "MISMATCH" in r.stdout or "FAULT" in r.stdout, r.stdout + r.stderr
