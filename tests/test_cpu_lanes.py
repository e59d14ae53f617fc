"""libhfx's kernels with 64-lane waves on the CPU (tests/cpuemu/libhfx_lanes.so:
every workgroup runs as its blockDim.x lanes, each a fiber, and ballots,
shuffles, wave barriers and __syncthreads_count complete only when every
lane of the wave or block that has not exited reaches them).  The one-lane
stand-in (libhfx_cpuemu.so) cannot show a mistake in code where lanes see
each other; these tests run that code as the GPU does:

* targeted checks (tests/cpuemu/lanes_unit.cpp): k_screen's ballot
  compaction, early stop and per-molecule segments against a direct loop
  over every quartet; k_screen_diag's wave-aggregated append; pick_resident's
  block vote; the stand-in's own collectives, and that it stops on a ballot
  reached by only some lanes of a wave;
* the emulation tests of the builds again with 64-lane waves: golden Fock
  matrices, every class's blocks against the oracle (FP64 and FP32), J only,
  mixed precision, chunked block classes (LDS rows of X, the per-wave X_cd
  staging of the digestion, pick_resident in every step), batch contexts
  and graph replays.  The SCF drivers' kernels (tree reductions over 256
  virtual blocks, one block barrier per level) are left to the one-lane
  stand-in and the GPU: with every lane a fiber they take minutes here.
"""
import ctypes
import os
import subprocess

import pytest

import tests.test_cpu_batch as B
import tests.test_cpu_emulation as E
from tests.conftest import LANES_LIB, ROOT
from tests.golden_io import CASES

HERE = os.path.join(ROOT, "tests", "cpuemu")


@pytest.fixture(scope="module")
def lanes_unit(tmp_path_factory):
    if not os.path.exists(LANES_LIB):
        from tests.cpuemu import build_cpuemu
        build_cpuemu.build(lanes=True)
    exe = str(tmp_path_factory.mktemp("lanes") / "lanes_unit")
    inc = ["-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "hfcode_amd", "csrc")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-DHFX_EMU_LANES"] + inc +
                   [os.path.join(HERE, "lanes_unit.cpp"), "-o", exe, LANES_LIB, "-Wl,-rpath," + HERE], check=True)
    return exe


def test_screen_compaction_and_votes(lanes_unit):
    r = subprocess.run([lanes_unit, "all"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and "all checks passed" in r.stdout
    assert r.stdout.count("ok k_screen ") == 10 and r.stdout.count("ok k_screen_diag") == 4
    assert "ok pick_resident" in r.stdout and "ok collectives" in r.stdout and "ok wave_sum" in r.stdout


def test_divergent_collective_is_reported(lanes_unit):
    r = subprocess.run([lanes_unit, "divergent"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=60)
    assert r.returncode != 0
    assert "lanes of one wave wait at different collectives" in r.stderr


@pytest.mark.parametrize("name", CASES)
def test_builds_match_golden(emu_lanes, name):
    E.test_builds_match_golden(emu_lanes, name)


@pytest.mark.parametrize("name", ["water_def2svp", "formaldehyde_631gmix"])
def test_class_blocks_match_oracle(emu_lanes, name):
    E.test_class_blocks_match_oracle(emu_lanes, name)


def test_fp32_class_blocks_match_oracle(emu_lanes):
    E.test_fp32_class_blocks_match_oracle(emu_lanes)


def test_j_only_matches_oracle(emu_lanes):
    E.test_j_only_matches_oracle(emu_lanes)


@pytest.mark.parametrize("name", ["water_def2svp", "formaldehyde_631gmix"])
def test_all_fp32_build_matches_fp64(emu_lanes, name):
    E.test_all_fp32_build_matches_fp64(emu_lanes, name)


@pytest.mark.parametrize("split,chunk", [(0.0, 256), (1e30, 0)])
def test_one_pass_digestion_equals_multi_pass(emu_lanes, split, chunk):
    E.test_one_pass_digestion_equals_multi_pass(emu_lanes, split, chunk)


def test_chunked_builds(emu_lanes):
    E.test_chunked_builds(emu_lanes)


@pytest.mark.parametrize("name", ["water_sto3g", "water_def2svp"])
def test_standalone_path_and_scf(emu_lanes, name):
    E.test_standalone_path_and_scf(emu_lanes, name)


@pytest.fixture(scope="module")
def lanes_batch():
    from hfcode_amd import fock
    if not os.path.exists(LANES_LIB):
        from tests.cpuemu import build_cpuemu
        build_cpuemu.build(lanes=True)
    saved = fock._lib
    fock._lib = fock.configure(ctypes.CDLL(LANES_LIB))
    st = B.make_setup(fock)
    yield st
    st["bb"].close()
    fock._lib = saved


def test_batch_builds_equal_alone(lanes_batch):
    B.test_batch_builds_equal_alone(lanes_batch, tol=1e-12)


def test_graph_replays_equal_uncaptured_builds(lanes_batch):
    B.test_graph_replays_equal_uncaptured_builds(lanes_batch)

