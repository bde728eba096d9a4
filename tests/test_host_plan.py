"""CPU tests of host-side sizing rules of the GPU path.

The split of k in the GEMM (matops_amd/csrc/gemm_plan.hpp) decides how many
partial tiles and chunk flags a launch uses; the entry points reserve and
clear what gemm_split_need reports. The sweep compiles the header on the
host and checks, for every M up to 40000 and a grid of N, K and CU counts,
that what the chosen instance uses fits in that figure.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "csrc", "gemm_plan_sweep.cpp")


@pytest.fixture(scope="module")
def sweep_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("plan") / "gemm_plan_sweep")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out, SRC], check=True)
    return out


def test_split_sizing_sweep(sweep_bin):
    r = subprocess.run([sweep_bin, "40000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
    assert int(r.stdout.split()[1]) > 1_000_000


def test_skinny_split_uses_more_flags_than_the_64x64_rule(sweep_bin):
    # M = 19585 (M mod 128 in [1, 64]), N = 16, K = 2048 on 256 CUs: the
    # skinny instance runs 154 row tiles x 8 chunks = 1232 flags, four more
    # than the 307 x 4 of the 64x64 rule; the reserved and cleared figure
    # must cover the larger one
    r = subprocess.run([sweep_bin, "c", "19585", "16", "2048", "256"], capture_output=True,
                       text=True, check=True)
    v = dict(kv.split("=") for kv in r.stdout.split())
    assert (int(v["S"]), int(v["S2"])) == (4, 8)
    assert int(v["flags_skinny"]) == 1232 and int(v["flags64"]) == 1228
    assert int(v["need_flags"]) >= 1232
