"""A C++ program written against parallax/stdpar.hpp (tests/cpp/funnel_dropin.cpp): on
the CPU every lookup misses and the funnels' ISO loops run; on the GPU the same binary's
fixed-function funnels resolve through the built-in table, copy_if through a descriptor
registered under its funnel key, and the work runs on the device."""
import os
import pathlib
import subprocess

import pytest

EXE = pathlib.Path(__file__).resolve().parent / "cpp" / "_bin" / "funnel_dropin"


def run(env_extra, *args):
    if not EXE.exists():
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    env = {k: v for k, v in os.environ.items() if k != "PARALLAX_FORCE_CPU"}
    env.update(env_extra)
    r = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=300, env=env)
    return r.returncode, r.stdout, r.stderr


def test_funnels_cpu_fallback():
    rc, out, err = run({"PARALLAX_FORCE_CPU": "1"})
    assert rc == 0, out + err
    assert "FAIL" not in out and "device_launches 0" in out


@pytest.mark.gpu
def test_funnels_on_gpu():
    rc, out, err = run({})
    assert rc == 0, out + err
    assert "FAIL" not in out
    launches = int(out.strip().splitlines()[-1].split()[-1])
    # reduce x2, scans x2, sort, unique, copy_if, partition, remove_if, count_if,
    # transform_reduce
    assert launches >= 11, out


@pytest.mark.gpu
def test_funnel_unique_beyond_the_arena_on_gpu():
    """parallax::unique of a sorted 2^28 uint64 vector (2 GiB, twice the default arena)
    runs on the device in place and gives std::unique's bytes."""
    rc, out, err = run({}, "big")
    assert rc == 0, out + err
    assert "PASS: unique 2^28 uint64 in place" in out
    assert int(out.strip().splitlines()[-1].split()[-1]) >= 1, out
